"""ctypes wrapper of the CPU oracle (oracle/_build/libmforacle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Every function is a thin call
into mf_oracle.c, which cites the reference code it restates.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmforacle.so")
if not os.path.exists(LIB):
    raise ImportError(f"{LIB} missing: run `make -C oracle`")
_o = C.CDLL(LIB)
_v = C.c_void_p


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _decl(name, res, args):
    f = getattr(_o, name)
    f.restype = res
    f.argtypes = args


_I = C.c_int
_ip = C.POINTER(C.c_int)
for name, res, args in [
    ("or_mtx_info", _I, [C.c_char_p, _ip, _ip, C.POINTER(C.c_int64), _ip]),
    ("or_mtx_read", _I, [C.c_char_p, _v, _v, _v]),
    ("or_tile_sizes", _I, [_I, _I, _v, _v, _ip, _ip, _ip, _ip, _ip]),
    ("or_tile_fill", _I, [_I, _I, _v, _v, _v, _v] + [_v] * 11),
    ("or_rowblkblock_count", _I, [_I, _v]),
    ("or_blockspmv_offsets", None, [_I, _I, _v, _v, _v, _v, _v, _v, _v]),
    ("or_encode", None, [_v, _v, _I, _I]),
    ("or_decode", None, [_v, _v, _I]),
    ("or_blockspmv_tiled", None, [_I, _I, _I, _I, _v, _v, _v, _v, _v, _v, _v, _v, _I]),
    ("or_blockspmv_nibble", None, [_I, _I, _I, _I, _v, _v, _v, _v, _v, _v, _v, _v]),
    ("or_csr_spmv", None, [_I, _v, _v, _v, _v, _v]),
    ("or_compact_count", _I, [_I, _I, _v, _v, _v, _v]),
    ("or_compact_fill", None, [_I, _I, _v, _v, _v, _v, _v, _v, _v, _v]),
    ("or_load_balance", _I, [_I, _I, _I, _v, _v, _I, _v, _v, _v, _ip, _ip]),
    ("or_shared_offsets", None, [_I, _v, _I, _v, _v, _v, _v]),
    ("or_cg", _I, [_I, _v, _v, _v, _v, _v, C.c_double, _I, C.POINTER(C.c_double)]),
    ("or_bicgstab", _I, [_I, _v, _v, _v, _v, _v, C.c_double, _I, C.POINTER(C.c_double)]),
    ("or_rel_residual", C.c_double, [_I, _v, _v, _v, _v, _v]),
    ("or_half_from_float", C.c_uint16, [C.c_float]),
    ("or_float_from_half", C.c_float, [C.c_uint16]),
    ("or_quant_fp16", None, [_v, _I, _v]),
    ("or_quant_int8", C.c_float, [_v, _I, _v]),
    ("or_rule_policy", None, [_I, _v, _v, _v]),
    ("or_mlp_forward", None, [_I, _v, _v, _v, _v, _v, _v, _v, _v]),
    ("or_mlp_forward16", None, [_I, _v, _v, _v, _v, _v, _v, _v, _v]),
    ("or_mlp_loss", C.c_double, [_I, _v, _v, _v, _v, _v, _v, _v, _v]),
    ("or_mlp_grad", C.c_double, [_I, _v, _v] + [_v] * 6 + [_v] * 6),
]:
    _decl(name, res, args)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# -- input ------------------------------------------------------------------
def read_mtx(path):
    m, n, s = C.c_int(), C.c_int(), C.c_int()
    nnz = C.c_int64()
    rc = _o.or_mtx_info(str(path).encode(), C.byref(m), C.byref(n), C.byref(nnz), C.byref(s))
    if rc:
        raise ValueError(f"or_mtx_info failed: {rc}")
    rp = np.zeros(m.value + 1, np.int64)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    va = np.zeros(max(nnz.value, 1), np.float64)
    rc = _o.or_mtx_read(str(path).encode(), _p(rp), _p(ci), _p(va))
    if rc:
        raise ValueError(f"or_mtx_read failed: {rc}")
    return m.value, n.value, rp, ci[: nnz.value], va[: nnz.value], s.value


def mtx_status(path):
    m, n, s = C.c_int(), C.c_int(), C.c_int()
    nnz = C.c_int64()
    return _o.or_mtx_info(str(path).encode(), C.byref(m), C.byref(n), C.byref(nnz), C.byref(s))


# -- tiles ------------------------------------------------------------------
def tile_create(rowA, colA, rowptr, colidx, val, val_low=None):
    rp, ci, va = _i32(rowptr), _i32(colidx), _f64(val)
    vl = np.ascontiguousarray(va.astype(np.float32) if val_low is None else val_low, np.float32)
    tm, tn, T, cs, cpl = (C.c_int() for _ in range(5))
    rc = _o.or_tile_sizes(rowA, colA, _p(rp), _p(ci), *(C.byref(v) for v in (tm, tn, T, cs, cpl)))
    if rc:
        raise ValueError("or_tile_sizes failed")
    T_ = T.value
    out = {
        "tile_ptr": np.zeros(tm.value + 1, np.int32), "tile_columnidx": np.zeros(max(T_, 1), np.int32),
        "tile_nnz": np.zeros(T_ + 1, np.int32), "blknnz": np.zeros(T_ + 1, np.int32),
        "blknnznnz": np.zeros(T_ + 1, np.uint8), "csr_offset": np.zeros(T_ + 1, np.int32),
        "csrptr_offset": np.zeros(T_ + 1, np.int32),
        "Blockcsr_Val": np.zeros(max(cs.value, 1), np.float64),
        "Blockcsr_Val_Low": np.zeros(max(cs.value, 1), np.float32),
        "Tile_csr_Col": np.zeros(max(cs.value, 1), np.uint8),
        "Blockcsr_Ptr": np.zeros(max(cpl.value, 1), np.uint8),
    }
    order = ["tile_ptr", "tile_columnidx", "tile_nnz", "blknnz", "blknnznnz", "csr_offset",
             "csrptr_offset", "Blockcsr_Val", "Blockcsr_Val_Low", "Tile_csr_Col", "Blockcsr_Ptr"]
    _o.or_tile_fill(rowA, colA, _p(rp), _p(ci), _p(va), _p(vl), *(_p(out[k]) for k in order))
    out["tile_columnidx"] = out["tile_columnidx"][:T_]
    for k in ("Blockcsr_Val", "Blockcsr_Val_Low", "Tile_csr_Col"):
        out[k] = out[k][: cs.value]
    out["Blockcsr_Ptr"] = out["Blockcsr_Ptr"][: cpl.value]
    out.update(tilem=tm.value, tilen=tn.value, tilenum=T_, csrsize=cs.value, csrptrlen=cpl.value,
               rowA=rowA, colA=colA)
    return out


def blockspmv_offsets(t):
    R = _o.or_rowblkblock_count(t["tilem"], _p(t["tile_ptr"]))
    T = t["tilenum"]
    p1, p2 = np.zeros(max(T, 1), np.int32), np.zeros(max(T, 1), np.int32)
    ri, cs, ce = np.zeros(max(R, 1), np.uint32), np.zeros(max(R, 1), np.int32), np.zeros(max(R, 1), np.int32)
    _o.or_blockspmv_offsets(t["rowA"], t["tilem"], _p(t["tile_ptr"]), _p(t["blknnz"]), _p(p1), _p(p2),
                            _p(ri), _p(cs), _p(ce))
    return {"rowblkblock": R, "ptroffset1": p1[:T], "ptroffset2": p2[:T],
            "blkcoostylerowidx": ri[:R], "colstart": cs[:R], "colstop": ce[:R]}


def encode(cols, offset=0):
    cols = np.ascontiguousarray(cols, np.uint8)
    out = np.zeros(offset + (len(cols) + 1) // 2 + 1, np.uint8)
    _o.or_encode(_p(cols), _p(out), len(cols), offset)
    return out[: offset + (len(cols) + 1) // 2]


def decode(packed, n):
    packed = np.ascontiguousarray(packed, np.uint8)
    out = np.zeros(max(n, 1), np.uint8)
    _o.or_decode(_p(packed), _p(out), n)
    return out[:n]


def xpad(t, x):
    xp = np.zeros(16 * max(t["tilen"], t["tilem"], 1), np.float64)
    xp[: len(x)] = x
    return xp


def blockspmv_tiled(t, x, nthreads=1):
    xp = xpad(t, x)
    y = np.zeros(16 * max(t["tilem"], 1), np.float64)
    _o.or_blockspmv_tiled(t["rowA"], t["colA"], t["tilem"], t["tilen"], _p(t["tile_ptr"]),
                          _p(t["tile_columnidx"]), _p(t["blknnz"]), _p(t["Blockcsr_Val"]),
                          _p(t["Tile_csr_Col"]), _p(t["Blockcsr_Ptr"]), _p(xp), _p(y), int(nthreads))
    return y[: t["rowA"]]


def blockspmv_nibble(t, x, cidx):
    xp = xpad(t, x)
    y = np.zeros(16 * max(t["tilem"], 1), np.float64)
    cidx = np.ascontiguousarray(cidx, np.uint8)
    _o.or_blockspmv_nibble(t["rowA"], t["colA"], t["tilem"], t["tilen"], _p(t["tile_ptr"]),
                           _p(t["tile_columnidx"]), _p(t["blknnz"]), _p(t["Blockcsr_Val"]), _p(cidx),
                           _p(t["Blockcsr_Ptr"]), _p(xp), _p(y))
    return y[: t["rowA"]]


def csr_spmv(rowptr, colidx, val, x, m=None):
    rp, ci, va, xx = _i32(rowptr), _i32(colidx), _f64(val), _f64(x)
    m = len(rp) - 1 if m is None else m
    y = np.zeros(max(m, 1), np.float64)
    _o.or_csr_spmv(m, _p(rp), _p(ci), _p(va), _p(xx), _p(y))
    return y[:m]


def compact(t):
    T = t["tilenum"]
    nzr = np.zeros(T + 1, np.int32)
    cnt = _o.or_compact_count(t["rowA"], t["tilem"], _p(t["tile_ptr"]), _p(t["blknnz"]),
                              _p(t["Blockcsr_Ptr"]), _p(nzr))
    rid, rpn = np.zeros(cnt + 1, np.uint8), np.zeros(cnt + 1, np.uint8)
    sig, tri = np.zeros(t["tilem"] + 1, np.int32), np.zeros(max(T, 1), np.int32)
    _o.or_compact_fill(t["rowA"], t["tilem"], _p(t["tile_ptr"]), _p(t["blknnz"]), _p(t["Blockcsr_Ptr"]),
                       _p(nzr), _p(rid), _p(rpn), _p(sig), _p(tri))
    return {"nonzero_row_new": nzr, "blockrowid_new": rid, "blockcsr_ptr_new": rpn,
            "block_signal": sig, "tile_rowidx": tri[:T]}


def load_balance(t, nnzR, warp_limit=6912):
    T = t["tilenum"]
    btp = np.zeros(T + 2, np.int32)
    ieb, reb = np.zeros(T + 1, np.int32), np.zeros(T + 1, np.int32)
    E, B = C.c_int(), C.c_int()
    idx = _o.or_load_balance(t["tilem"], T, nnzR, _p(t["tile_ptr"]), _p(t["blknnz"]), warp_limit,
                             _p(btp), _p(ieb), _p(reb), C.byref(E), C.byref(B))
    if idx < 0:
        return None
    return {"nwarps": idx, "balance_tile_ptr": btp[: idx + 1], "index_each_block": ieb,
            "row_each_block": reb, "each_block_nnz": E.value, "block_per_warp": B.value}


def shared_offsets(t, bal):
    """src/main-cg-mixed.cu:2681-2708 restated: shared-queue offsets of a
    load_balance() result."""
    T, nw = t["tilenum"], int(bal["nwarps"])
    off = np.zeros(T + 1, np.int32)
    end = np.zeros(nw + 1, np.int32)
    _o.or_shared_offsets(T, _p(_i32(t["blknnz"])), nw, _p(_i32(bal["balance_tile_ptr"])),
                         _p(_i32(bal["index_each_block"])), _p(off), _p(end))
    return {"non_each_block_offset": off, "balance_tile_ptr_shared_end": end}


# -- solvers ------------------------------------------------------------------
def cg(rowptr, colidx, val, b, x0=None, tol=1e-10, maxiter=10000):
    rp, ci, va, bb = _i32(rowptr), _i32(colidx), _f64(val), _f64(b)
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    rel = C.c_double()
    it = _o.or_cg(n, _p(rp), _p(ci), _p(va), _p(bb), _p(x), tol, maxiter, C.byref(rel))
    return x, it, rel.value


def bicgstab(rowptr, colidx, val, b, x0=None, tol=1e-10, maxiter=10000):
    rp, ci, va, bb = _i32(rowptr), _i32(colidx), _f64(val), _f64(b)
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    rel = C.c_double()
    it = _o.or_bicgstab(n, _p(rp), _p(ci), _p(va), _p(bb), _p(x), tol, maxiter, C.byref(rel))
    return x, it, rel.value


def rel_residual(rowptr, colidx, val, b, x):
    rp, ci, va, bb, xx = _i32(rowptr), _i32(colidx), _f64(val), _f64(b), _f64(x)
    return _o.or_rel_residual(len(bb), _p(rp), _p(ci), _p(va), _p(bb), _p(xx))


# -- payloads / policy ----------------------------------------------------------
def half_from_float(f):
    return _o.or_half_from_float(float(f))


def float_from_half(h):
    return _o.or_float_from_half(int(h))


def quant_fp16(v):
    v = _f64(v)
    out = np.zeros(max(len(v), 1), np.uint16)
    _o.or_quant_fp16(_p(v), len(v), _p(out))
    return out[: len(v)]


def quant_int8(v):
    v = _f64(v)
    out = np.zeros(max(len(v), 1), np.int8)
    s = _o.or_quant_int8(_p(v), len(v), _p(out))
    return out[: len(v)], s


def rule_policy(p_new, p_old):
    a, b = _f64(p_new), _f64(p_old)
    nseg = len(a) // 16
    out = np.zeros(max(nseg, 1), np.uint8)
    _o.or_rule_policy(nseg, _p(a), _p(b), _p(out))
    return out[:nseg]


# -- MLP ------------------------------------------------------------------------
def mlp_split(params):
    p = _f64(params)
    sizes = [32 * 8, 32, 32 * 32, 32, 4 * 32, 4]
    out, o = [], 0
    for s in sizes:
        out.append(np.ascontiguousarray(p[o:o + s]))
        o += s
    return out


def mlp_forward(params, feat):
    w = mlp_split(params)
    f = _f64(feat).ravel()
    n = len(f) // 8
    z = np.zeros(max(n, 1) * 4)
    _o.or_mlp_forward(n, _p(f), *(_p(a) for a in w), _p(z))
    return z[: 4 * n].reshape(n, 4)


def mlp_forward16(params, feat):
    """Forward pass with the device's FP16 rounding of parameters, features
    and hidden activations (the parameters are rounded here)."""
    p16 = np.asarray(params, dtype=np.float32).astype(np.float16).astype(np.float64)
    w = mlp_split(p16)
    f = _f64(feat).ravel()
    n = len(f) // 8
    z = np.zeros(max(n, 1) * 4)
    _o.or_mlp_forward16(n, _p(f), *(_p(a) for a in w), _p(z))
    return z[: 4 * n].reshape(n, 4)


def mlp_loss(params, feat, label):
    w = mlp_split(params)
    f, lab = _f64(feat), _i32(label)
    return _o.or_mlp_loss(len(lab), _p(f), _p(lab), *(_p(a) for a in w))


def mlp_grad(params, feat, label):
    w = mlp_split(params)
    f, lab = _f64(feat), _i32(label)
    g = [np.zeros_like(a) for a in w]
    L = _o.or_mlp_grad(len(lab), _p(f), _p(lab), *(_p(a) for a in w), *(_p(a) for a in g))
    return L, np.concatenate(g)
