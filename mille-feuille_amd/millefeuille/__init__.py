"""millefeuille -- MI355X tiled mixed-precision SpMV / CG / BiCGSTAB.

Python view of libmillefeuille_hip.so (C ABI: include/millefeuille.h).
Importing this package loads the native library and fails if it is missing.
"""
from . import _lib  # noqa: F401  (loads the native library, raises if absent)
from ._lib import (FORM_AUTO, FORM_MEMORY, FORM_RESIDENT, MLP_NPARAM, POLICY_ERRBOUND,
                   POLICY_FIXED, POLICY_FP64, POLICY_LEARNED, POLICY_RULE, PREC_FP16, PREC_FP32,
                   PREC_FP64, PREC_INT8, PREC_SKIP, LIB_PATH, MfError)
from .core import TiledMatrix, batch_cg, read_mtx, solve_opts, spmv_bytes
from .hip import DeviceArray, Timer, device_count, synchronize

__all__ = ["TiledMatrix", "batch_cg", "read_mtx", "solve_opts", "spmv_bytes", "DeviceArray", "Timer",
           "device_count", "synchronize", "MfError", "LIB_PATH", "MLP_NPARAM",
           "PREC_FP64", "PREC_FP32", "PREC_FP16", "PREC_INT8", "PREC_SKIP",
           "POLICY_FP64", "POLICY_RULE", "POLICY_FIXED", "POLICY_LEARNED", "POLICY_ERRBOUND",
           "FORM_AUTO", "FORM_MEMORY", "FORM_RESIDENT"]
