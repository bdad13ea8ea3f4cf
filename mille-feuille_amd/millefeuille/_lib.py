"""ctypes binding of libmillefeuille_hip.so (the C ABI in include/millefeuille.h).

The library is loaded from this directory; there is no fallback: if it is
missing the import raises.  Device buffers are managed through the HIP
runtime the library itself links (libamdhip64.so.7), see hip.py.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmillefeuille_hip.so")
# the reference-named entry points with mixed-precision defaults (the drop-in
# for the mixed executables); loaded on demand by its users
MIXED_LIB_PATH = os.path.join(_HERE, "libmillefeuille_mixed.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C mille-feuille_amd` "
        "(or __graft_entry__.build()); there is no CPU fallback")

lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)

MF_OK = 0
MF_NOT_CONVERGED = 1
PREC_FP64, PREC_FP32, PREC_FP16, PREC_INT8, PREC_SKIP = range(5)
POLICY_FP64, POLICY_RULE, POLICY_FIXED, POLICY_LEARNED, POLICY_ERRBOUND = range(5)
SPMV_AUTO, SPMV_STREAM, SPMV_ROWGROUP = -1, 0, 1
FORM_AUTO, FORM_MEMORY, FORM_RESIDENT = 0, 1, 2
MLP_NPARAM = 32 * 8 + 32 + 32 * 32 + 32 + 4 * 32 + 4


class MfCsr(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int64),
                ("rowptr", C.POINTER(C.c_int)), ("colidx", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double)), ("is_sym", C.c_int)]


class MfTileArrays(C.Structure):
    _fields_ = [("tile_ptr", C.POINTER(C.c_int)), ("tile_columnidx", C.POINTER(C.c_int)),
                ("tile_nnz", C.POINTER(C.c_int)), ("blknnz", C.POINTER(C.c_int)),
                ("csr_offset", C.POINTER(C.c_int)), ("csrptr_offset", C.POINTER(C.c_int)),
                ("blknnznnz", C.POINTER(C.c_ubyte)), ("Tile_csr_Col", C.POINTER(C.c_ubyte)),
                ("Blockcsr_Ptr", C.POINTER(C.c_ubyte)), ("Blockcsr_Val", C.POINTER(C.c_double)),
                ("Blockcsr_Val_Low", C.POINTER(C.c_float))]


class MfSolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int), ("policy", C.c_int),
                ("check_interval", C.c_int), ("refresh_ratio", C.c_double),
                ("max_refresh", C.c_int), ("refresh_every", C.c_int), ("blocks", C.c_int),
                ("fixed_iters", C.c_int), ("form", C.c_int), ("batch_waves", C.c_int),
                ("stream", C.c_void_p)]


class MfReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refreshes", C.c_int), ("fell_back", C.c_int),
                ("status", C.c_int), ("rel_residual", C.c_double),
                ("rel_residual_recursive", C.c_double), ("time_ms", C.c_double),
                ("prec_hist", C.c_longlong * 5), ("grid_blocks", C.c_int),
                ("form_used", C.c_int), ("policy_used", C.c_int)]

    def as_dict(self):
        return {"iterations": self.iterations, "refreshes": self.refreshes,
                "fell_back": self.fell_back, "status": self.status,
                "rel_residual": self.rel_residual,
                "rel_residual_recursive": self.rel_residual_recursive,
                "time_ms": self.time_ms, "prec_hist": list(self.prec_hist),
                "grid_blocks": self.grid_blocks, "form_used": self.form_used,
                "policy_used": self.policy_used}


class MfTuneReport(C.Structure):
    _fields_ = [("ncand", C.c_int), ("cand", C.c_int * 8), ("us", C.c_float * 8), ("best", C.c_int)]

    def as_dict(self):
        n = self.ncand
        return {"cand": list(self.cand[:n]), "us": [float(v) for v in self.us[:n]], "best": self.best}


_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


# Every symbol declared in include/millefeuille.h (the ABI test checks this
# list against the header).
SIGNATURES = {
    "mf_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(MfCsr)]),
    "mf_read_cbd": (C.c_int, [C.c_char_p, C.POINTER(MfCsr)]),
    "mf_write_cbd": (C.c_int, [C.c_char_p, C.POINTER(MfCsr)]),
    "mf_csr_free": (None, [C.POINTER(MfCsr)]),
    "mf_tile_create": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "mf_tile_free": (None, [_vp]),
    "mf_tile_info": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip, _ip]),
    "mf_tile_arrays_get": (C.c_int, [_vp, C.POINTER(MfTileArrays)]),
    "mf_tile_rowblkblock": (C.c_int, [_vp, _ip]),
    "mf_tile_offsets": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_tile_compact_count": (C.c_int, [_vp, _vp, _ip]),
    "mf_tile_compact_fill": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_tile_shared_offsets": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "mf_tile_balance": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _ip]),
    "mf_encode_nibbles": (None, [_vp, _vp, C.c_int, C.c_int]),
    "mf_decode_nibbles": (None, [_vp, _vp, C.c_int]),
    "mf_tile_set_precision": (C.c_int, [_vp, _vp, C.c_int]),
    "mf_tile_get_precision": (C.c_int, [_vp, _vp]),
    "mf_tile_get_scales": (C.c_int, [_vp, _vp, _vp]),
    "mf_tile_get_quant_error": (C.c_int, [_vp, _vp]),
    "mf_spmv": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "mf_stream_copy": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
    "mf_tile_set_schedule": (C.c_int, [_vp, C.c_int]),
    "mf_tile_set_spmv_kernel": (C.c_int, [_vp, C.c_int]),
    "mf_solve_opts_default": (None, [C.POINTER(MfSolveOpts)]),
    "mf_cg_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(MfSolveOpts), C.POINTER(MfReport)]),
    "mf_cg_solve_noalloc": (C.c_int, [_vp, _vp, _vp, C.POINTER(MfSolveOpts), C.POINTER(MfReport)]),
    "mf_bicgstab_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(MfSolveOpts), C.POINTER(MfReport)]),
    "mf_cg_solve_multikernel": (C.c_int, [_vp, _vp, _vp, C.POINTER(MfSolveOpts), C.POINTER(MfReport)]),
    "mf_bicgstab_solve_multikernel": (C.c_int, [_vp, _vp, _vp, C.POINTER(MfSolveOpts), C.POINTER(MfReport)]),
    "cg_solve_reduce": (None, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int]),
    "cg_solve_sync": (None, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int]),
    "cg_solve_inc": (None, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int]),
    "cg_solve_mixed": (None, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int]),
    "bicgstab": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_char_p, C.c_int]),
    "bicgstab_sync": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_char_p, C.c_int]),
    "bicgstab_inc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_char_p, C.c_int]),
    "bicgstab_mixed": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_char_p, C.c_int]),
    "read_Dmatrix_32": (None, [_ip, _ip, _ip, _vp, _vp, _vp, _ip, C.c_char_p]),
    "mf_mlp_forward": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "mf_mlp_grad": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    "mf_mlp_train": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint,
                               C.POINTER(C.c_float), _vp]),
    "mf_policy_set_mlp": (C.c_int, [_vp, _vp]),
    "mf_policy_features": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp]),
    "mf_policy_labels": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp]),
    "mf_tune_spmv": (C.c_int, [_vp, C.c_int, C.POINTER(MfTuneReport)]),
    "mf_tune_batch": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(MfSolveOpts), C.c_int,
                                C.POINTER(MfTuneReport)]),
    "mf_batch_cg_solve": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(MfSolveOpts), _vp,
                                    _vp, _vp, C.POINTER(C.c_double)]),
    "mf_last_error": (C.c_char_p, []),
    "mf_version": (C.c_char_p, []),
}

for _n, (_r, _a) in SIGNATURES.items():
    _sig(_n, _r, _a)


class MfError(RuntimeError):
    def __init__(self, fn, code):
        msg = lib.mf_last_error().decode(errors="replace")
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def check(fn, code, allow=(MF_OK,)):
    if code not in allow:
        raise MfError(fn, code)
    return code
