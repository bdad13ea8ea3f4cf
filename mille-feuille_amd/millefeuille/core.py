"""Python view of the C ABI: a tiled matrix handle, SpMV and the solvers.

Mirrors the reference's operator surface for this path (Tile_create, the
blockspmv_cpu offsets, the compaction and balancer arrays, the CG/BiCGSTAB
entry points) with numpy arrays for host data and DeviceArray for HBM data.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib
from ._lib import MfCsr, MfReport, MfSolveOpts, MfTileArrays, check, lib
from .hip import DeviceArray

_POLICIES = {"fp64": _lib.POLICY_FP64, "rule": _lib.POLICY_RULE, "fixed": _lib.POLICY_FIXED,
             "learned": _lib.POLICY_LEARNED, "errbound": _lib.POLICY_ERRBOUND}


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def read_mtx(path):
    """Matrix Market -> (m, n, rowptr, colidx, val, is_sym), file order."""
    a = MfCsr()
    check("mf_read_mtx", lib.mf_read_mtx(str(path).encode(), C.byref(a)))
    try:
        rp = np.ctypeslib.as_array(a.rowptr, (a.m + 1,)).copy()
        ci = np.ctypeslib.as_array(a.colidx, (max(a.nnz, 1),))[: a.nnz].copy()
        va = np.ctypeslib.as_array(a.val, (max(a.nnz, 1),))[: a.nnz].copy()
        return a.m, a.n, rp, ci, va, a.is_sym
    finally:
        lib.mf_csr_free(C.byref(a))


_FORMS = {"auto": _lib.FORM_AUTO, "memory": _lib.FORM_MEMORY, "resident": _lib.FORM_RESIDENT}


def solve_opts(tol=1e-10, maxiter=10000, policy="fp64", check_interval=8, refresh_ratio=0.1,
               max_refresh=64, refresh_every=64, blocks=0, fixed_iters=0, form="auto", batch_waves=0):
    o = MfSolveOpts()
    lib.mf_solve_opts_default(C.byref(o))
    o.tol = tol
    o.maxiter = maxiter
    o.policy = _POLICIES[policy] if isinstance(policy, str) else int(policy)
    o.check_interval = check_interval
    o.refresh_ratio = refresh_ratio
    o.max_refresh = max_refresh
    o.refresh_every = refresh_every
    o.blocks = blocks
    o.fixed_iters = fixed_iters
    o.form = _FORMS[form] if isinstance(form, str) else int(form)
    o.batch_waves = int(batch_waves)
    return o


class TiledMatrix:
    """Handle on the 16x16 tiled matrix (host reference arrays + HBM layout)."""

    def __init__(self, rowptr, colidx, val, rowA=None, colA=None, val_low=None, device=0):
        self._rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        self._ci = np.ascontiguousarray(colidx, dtype=np.int32)
        self._va = np.ascontiguousarray(val, dtype=np.float64)
        self._vl = None if val_low is None else np.ascontiguousarray(val_low, dtype=np.float32)
        m = len(self._rp) - 1
        self.rowA = m if rowA is None else int(rowA)
        self.colA = m if colA is None else int(colA)
        h = C.c_void_p()
        check("mf_tile_create", lib.mf_tile_create(self.rowA, self.colA, _p(self._rp), _p(self._ci),
                                                   _p(self._va), _p(self._vl), device, C.byref(h)))
        self.h = h
        vals = [C.c_int() for _ in range(6)]
        check("mf_tile_info", lib.mf_tile_info(h, *[C.byref(v) for v in vals]))
        self.tilem, self.tilen, self.tilenum, self.csrsize, self.csrptrlen, self.npad = [v.value for v in vals]
        self._spmv_kernel = _lib.SPMV_STREAM

    def free(self):
        if getattr(self, "h", None):
            lib.mf_tile_free(self.h)
            self.h = None

    def __del__(self):
        # nothing is released through HIP during interpreter shutdown (the
        # runtime's exit handlers may have run; process exit frees it all)
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass

    # -- reference arrays -------------------------------------------------
    def arrays(self):
        a = MfTileArrays()
        check("mf_tile_arrays_get", lib.mf_tile_arrays_get(self.h, C.byref(a)))
        T, Z = self.tilenum, self.csrsize

        def cp(ptr, n):
            return np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].copy()

        return {"tile_ptr": cp(a.tile_ptr, self.tilem + 1),
                "tile_columnidx": cp(a.tile_columnidx, T),
                "tile_nnz": cp(a.tile_nnz, T + 1), "blknnz": cp(a.blknnz, T + 1),
                "csr_offset": cp(a.csr_offset, T + 1), "csrptr_offset": cp(a.csrptr_offset, T + 1),
                "blknnznnz": cp(a.blknnznnz, T + 1), "Tile_csr_Col": cp(a.Tile_csr_Col, Z),
                "Blockcsr_Ptr": cp(a.Blockcsr_Ptr, self.csrptrlen),
                "Blockcsr_Val": cp(a.Blockcsr_Val, Z), "Blockcsr_Val_Low": cp(a.Blockcsr_Val_Low, Z)}

    def offsets(self):
        rb = C.c_int()
        check("mf_tile_rowblkblock", lib.mf_tile_rowblkblock(self.h, C.byref(rb)))
        T, R = self.tilenum, rb.value
        p1 = np.zeros(max(T, 1), np.int32)
        p2 = np.zeros(max(T, 1), np.int32)
        ri = np.zeros(max(R, 1), np.uint32)
        cs = np.zeros(max(R, 1), np.int32)
        ce = np.zeros(max(R, 1), np.int32)
        check("mf_tile_offsets", lib.mf_tile_offsets(self.h, _p(p1), _p(p2), _p(ri), _p(cs), _p(ce)))
        return {"ptroffset1": p1[:T], "ptroffset2": p2[:T], "blkcoostylerowidx": ri[:R],
                "colstart": cs[:R], "colstop": ce[:R]}

    def compact(self):
        T = self.tilenum
        nzr = np.zeros(T + 1, np.int32)
        cnt = C.c_int()
        check("mf_tile_compact_count", lib.mf_tile_compact_count(self.h, _p(nzr), C.byref(cnt)))
        c = cnt.value
        rid = np.zeros(c + 1, np.uint8)
        rpn = np.zeros(c + 1, np.uint8)
        sig = np.zeros(self.tilem + 1, np.int32)
        tri = np.zeros(max(T, 1), np.int32)
        check("mf_tile_compact_fill", lib.mf_tile_compact_fill(self.h, _p(nzr), _p(rid), _p(rpn), _p(sig), _p(tri)))
        return {"nonzero_row_new": nzr, "blockrowid_new": rid, "blockcsr_ptr_new": rpn,
                "block_signal": sig, "tile_rowidx": tri[:T]}

    def balance(self, warp_limit=6912):
        T = self.tilenum
        btp = np.zeros(T + 2, np.int32)
        ieb = np.zeros(T + 1, np.int32)
        reb = np.zeros(T + 1, np.int32)
        nw = C.c_int()
        check("mf_tile_balance", lib.mf_tile_balance(self.h, warp_limit, _p(btp), _p(ieb), _p(reb), C.byref(nw)))
        return {"nwarps": nw.value, "balance_tile_ptr": btp[: nw.value + 1],
                "index_each_block": ieb, "row_each_block": reb}

    def shared_offsets(self, bal):
        """The reference's shared-queue offsets (mf_tile_shared_offsets) of a
        balance() result."""
        nw = int(bal["nwarps"])
        btp = np.ascontiguousarray(bal["balance_tile_ptr"], np.int32)
        ieb = np.ascontiguousarray(bal["index_each_block"], np.int32)
        off = np.zeros(self.tilenum + 1, np.int32)
        end = np.zeros(nw + 1, np.int32)
        check("mf_tile_shared_offsets", lib.mf_tile_shared_offsets(self.h, nw, _p(btp), _p(ieb), _p(off), _p(end)))
        return {"non_each_block_offset": off, "balance_tile_ptr_shared_end": end}

    # -- precision ---------------------------------------------------------
    def set_precision(self, levels=None, uniform=0):
        lv = None if levels is None else np.ascontiguousarray(levels, dtype=np.uint8)
        if lv is not None and len(lv) != self.tilenum:
            raise ValueError("one level per tile expected")
        check("mf_tile_set_precision", lib.mf_tile_set_precision(self.h, _p(lv), int(uniform)))

    def set_schedule(self, spmv_wave_nnz):
        check("mf_tile_set_schedule", lib.mf_tile_set_schedule(self.h, int(spmv_wave_nnz)))
        if int(spmv_wave_nnz) > 0:
            self._wave_nnz = int(spmv_wave_nnz)

    def tune_spmv(self, mode=0):
        """Autotunes the stand-alone SpMV (mf_tune_spmv) for `mode` at the
        handle's current levels and keeps the fastest schedule; returns the
        candidates' timings."""
        rep = _lib.MfTuneReport()
        check("mf_tune_spmv", lib.mf_tune_spmv(self.h, int(mode), C.byref(rep)))
        d = rep.as_dict()
        self._spmv_kernel = _lib.SPMV_ROWGROUP if d["best"] == 0 else _lib.SPMV_STREAM
        if d["best"] > 0:
            self._wave_nnz = int(d["best"])
        return d

    def spmv_kernel_name(self):
        """Device kernel that mf_spmv launches for this handle (mode 0): the
        row-group kernel, or the streamed one, instantiated with piece
        bookkeeping when a tile-row holds more than 4*max(wave_nnz, 4096)
        padded payload slots (the library's heavy tile-row rule)."""
        if self._spmv_kernel == _lib.SPMV_ROWGROUP:
            return "mf::k_spmv_rg<0>"
        a = self.arrays()
        slots = (np.diff(a["csr_offset"]).astype(np.int64) + 3) // 4 * 4
        pre = np.concatenate([[0], np.cumsum(slots)])
        row_slots = np.diff(pre[np.asarray(a["tile_ptr"], np.int64)])
        heavy = bool((row_slots > 4 * max(getattr(self, "_wave_nnz", 1024), 4096)).any())
        return f"mf::k_spmv<0, {'true' if heavy else 'false'}>"

    def set_spmv_kernel(self, kind):
        """'auto', 'stream' or 'rowgroup' (or the MF_SPMV_* value)."""
        k = {"auto": _lib.SPMV_AUTO, "stream": _lib.SPMV_STREAM, "rowgroup": _lib.SPMV_ROWGROUP}
        kind = k.get(kind, kind)
        check("mf_tile_set_spmv_kernel", lib.mf_tile_set_spmv_kernel(self.h, kind))
        self._spmv_kernel = _lib.SPMV_STREAM if kind == _lib.SPMV_AUTO else kind

    def get_precision(self):
        out = np.zeros(max(self.tilenum, 1), np.uint8)
        check("mf_tile_get_precision", lib.mf_tile_get_precision(self.h, _p(out)))
        return out[: self.tilenum]

    def scales(self):
        s16 = np.zeros(max(self.tilenum, 1), np.float32)
        s8 = np.zeros(max(self.tilenum, 1), np.float32)
        check("mf_tile_get_scales", lib.mf_tile_get_scales(self.h, _p(s16), _p(s8)))
        return s16[: self.tilenum], s8[: self.tilenum]

    def quant_error(self):
        e = np.zeros(max(3 * self.tilenum, 1), np.float64)
        check("mf_tile_get_quant_error", lib.mf_tile_get_quant_error(self.h, _p(e)))
        return e[: 3 * self.tilenum].reshape(3, self.tilenum)

    # -- compute -----------------------------------------------------------
    def vector(self, host=None, dtype=np.float64):
        """Device vector of npad entries (zero padded)."""
        if host is None:
            d = DeviceArray(self.npad, dtype)
            d.zero()
            return d
        return DeviceArray.from_host(np.asarray(host, dtype=dtype), length=self.npad)

    def spmv(self, d_x, d_y, mode=0, d_x32=None, stream=None):
        check("mf_spmv", lib.mf_spmv(self.h, d_x.ptr, None if d_x32 is None else d_x32.ptr,
                                     d_y.ptr, mode, stream))

    def _solve(self, fn, d_b, d_x, opts, allow):
        rep = MfReport()
        code = fn(self.h, d_b.ptr, d_x.ptr, C.byref(opts), C.byref(rep))
        check(fn.__name__, code, allow=allow)
        return rep.as_dict()

    def cg(self, d_b, d_x, allow_not_converged=False, **kw):
        allow = (0, 1) if allow_not_converged else (0,)
        return self._solve(lib.mf_cg_solve, d_b, d_x, solve_opts(**kw), allow)

    def cg_noalloc(self, d_b, d_x, allow_not_converged=False, **kw):
        """mf_cg_solve_noalloc: the same solve with no device allocation (the
        library's fixed scratch array; systems up to about 150,000 rows)."""
        allow = (0, 1) if allow_not_converged else (0,)
        return self._solve(lib.mf_cg_solve_noalloc, d_b, d_x, solve_opts(**kw), allow)

    def bicgstab(self, d_b, d_x, allow_not_converged=False, **kw):
        allow = (0, 1) if allow_not_converged else (0,)
        return self._solve(lib.mf_bicgstab_solve, d_b, d_x, solve_opts(**kw), allow)

    def cg_multikernel(self, d_b, d_x, allow_not_converged=False, **kw):
        allow = (0, 1) if allow_not_converged else (0,)
        return self._solve(lib.mf_cg_solve_multikernel, d_b, d_x, solve_opts(**kw), allow)

    def bicgstab_multikernel(self, d_b, d_x, allow_not_converged=False, **kw):
        allow = (0, 1) if allow_not_converged else (0,)
        return self._solve(lib.mf_bicgstab_solve_multikernel, d_b, d_x, solve_opts(**kw), allow)


def spmv_bytes(tm: "TiledMatrix", levels=None, x_bytes=8):
    """Algorithmic bytes of one SpMV: sum_t nnz_t*(s(p_t)+1) + 24*T
    + 4*T_scaled + n*s_x + 8*n (value + 1-byte local column per entry; per
    tile 16 B row pointers, 4 B tile column, 4 B payload offset; a 4 B scale
    per FP16/INT8 tile; the x read and the y write)."""
    a = tm.arrays()
    nz = np.diff(a["csr_offset"]).astype(np.int64)
    size = np.array([8, 4, 2, 1, 0], np.int64)
    if levels is None:
        levels = np.zeros(tm.tilenum, np.uint8)
    lv = np.asarray(levels)
    payload = int(np.sum(nz * np.where(lv == 4, 0, size[lv] + 1)))
    scaled = int(np.sum((lv == 2) | (lv == 3)))
    active = int(np.sum(lv != 4))
    n = tm.rowA
    return payload + 24 * active + 4 * scaled + n * x_bytes + 8 * n


def batch_cg(systems, bs, nrhs=1, x0=None, allow_not_converged=False, tune_iters=0, **kw):
    """Solves every system in `systems` (TiledMatrix handles on one device)
    for `nrhs` right-hand sides each, in one persistent launch captured in a
    hipGraph (mf_batch_cg_solve).  bs[i] is an (nrhs, rowA_i) or (rowA_i,)
    array.  tune_iters > 0 first picks the waves per item with mf_tune_batch
    on that many iterations (unless batch_waves is given).  Returns (xs,
    report) with xs[i] of shape (nrhs, rowA_i) and a report holding
    iterations, relative residuals, per-system and total ms (and the tuner's
    timings under "tune")."""
    opts = solve_opts(**kw)
    nsys = len(systems)
    db, dx = [], []
    for tm, b in zip(systems, bs):
        b = np.asarray(b, np.float64).reshape(nrhs, -1)
        hb = np.zeros((nrhs, tm.npad))
        hb[:, : b.shape[1]] = b
        db.append(DeviceArray.from_host(hb.ravel()))
        hx = np.zeros((nrhs, tm.npad))
        if x0 is not None:
            x = np.asarray(x0[len(dx)], np.float64).reshape(nrhs, -1)
            hx[:, : x.shape[1]] = x
        dx.append(DeviceArray.from_host(hx.ravel()))
    hs = (C.c_void_p * nsys)(*[tm.h.value for tm in systems])
    pb = (C.c_void_p * nsys)(*[d.ptr for d in db])
    px = (C.c_void_p * nsys)(*[d.ptr for d in dx])
    tune = None
    if tune_iters > 0 and not kw.get("batch_waves"):
        rep = _lib.MfTuneReport()
        check("mf_tune_batch", lib.mf_tune_batch(hs, nsys, nrhs, pb, px, C.byref(opts), int(tune_iters),
                                                 C.byref(rep)))
        tune = rep.as_dict()
        opts.batch_waves = tune["best"]
    its = np.zeros(nsys * nrhs, np.int32)
    rel = np.zeros(nsys * nrhs, np.float64)
    per = np.zeros(nsys, np.float64)
    tot = C.c_double()
    code = lib.mf_batch_cg_solve(hs, nsys, nrhs, pb, px, C.byref(opts), _p(its), _p(rel), _p(per),
                                 C.byref(tot))
    check("mf_batch_cg_solve", code, allow=(0, 1) if allow_not_converged else (0,))
    xs = [d.get().reshape(nrhs, tm.npad)[:, : tm.rowA] for d, tm in zip(dx, systems)]
    return xs, {"status": code, "iterations": its.reshape(nsys, nrhs),
                "rel_residual": rel.reshape(nsys, nrhs), "per_system_ms": per,
                "total_ms": tot.value, "batch_waves": opts.batch_waves or 8, "tune": tune}
