"""ctypes binding of libdmc_gpu.so (the C-ABI declared in include/dmc_gpu.h).

There is no CPU fallback: if the library is missing or no HIP device is present, every
entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libdmc_gpu.so")
MAXORB = 16

DMC_FP64, DMC_MIXED = 0, 1
DMC_ALLPARTICLE, DMC_SINGLEPARTICLE, DMC_VMC, DMC_METROPOLIS = 0, 1, 2, 3
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "population overflow",
          4: "set-up failure", 5: "no HIP device"}


class DMCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class dmc_params(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("npart", "norb", "ntab", "fn", "units", "nwalk",
                                         "capacity", "precision", "movetype", "krefresh")] + [
        (k, C.c_double) for k in ("rho", "hbar", "mass", "dt", "etrial", "dcut", "wtcut",
                                  "popup", "popdn", "bbeta", "dbeta", "bfact", "v0", "amu",
                                  "cjas1", "cjas2", "djas", "v0op", "v0eq")] + [
        ("vorb", C.c_double * MAXORB)]


_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# (name, argtypes) of every exported entry point of include/dmc_gpu.h
SIGNATURES = {
    "dmc_create": [C.POINTER(dmc_params), C.POINTER(_vp)],
    "dmc_destroy": [_vp],
    "dmc_info": [_vp, _f64],
    "dmc_last_error": [],
    "dmc_set_tables": [_vp] + [_f64] * 7,
    "dmc_set_params": [_vp, _f64, C.c_double, C.c_double],
    "dmc_load_walkers": [_vp, C.c_int32, _f64, _i64],
    "dmc_init_walkers": [_vp, C.c_int64, C.c_int32],
    "dmc_read_walkers": [_vp, C.c_int32, C.POINTER(C.c_int32), _vp, _vp, _vp, _vp],
    "dmc_population": [_vp, C.POINTER(C.c_int32)],
    "dmc_eval": [_vp, C.c_int32, _f64, _f64, _i32, _f64, _f64, _f64],
    "dmc_ratio1": [_vp, C.c_int32, _f64, _i32, _f64, _f64],
    "dmc_step": [_vp, C.c_int32, _i32],
    "dmc_propagate": [_vp, C.c_int32],
    "dmc_branch": [_vp, C.c_double, C.POINTER(C.c_int32)],
    "dmc_run": [_vp, C.c_int32, C.POINTER(C.c_int32)],
    "dmc_estimators": [_vp, _vp, C.c_int32],
    "dmc_wtisto": [_vp, _i64],
    "dmc_acceptance": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "dmc_adjoint": [_vp, C.c_int32, _f64, _f64, _vp],
    "dmc_opt_accumulate": [_vp],
    "dmc_opt_accumulate_noalloc": [_vp],
    "dmc_opt_sums": [_vp, _f64, C.c_int32],
    "dmc_gofr": [_vp, C.c_int32, _f64, _f64, _f64, _f64, _f64],
    "dmc_sofq_kvec": [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _vp, _vp],
    "dmc_sofq": [_vp, C.c_int32, _f64, _f64, _f64],
    "dmc_nofk_shells": [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _vp, _vp],
    "dmc_nofk": [_vp, C.c_int32, _f64, _f64, _f64],
    "dmc_move_trace": [_vp, C.c_int32],
    "dmc_read_move_trace": [_vp, C.c_int32, C.POINTER(C.c_int32), _vp],
    "dmc_kernel_timing": [_vp, C.c_int32],
    "dmc_kernel_times": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32],
    "dmc_sync": [_vp],
    "dmc_batch_create": [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)],
    "dmc_batch_destroy": [_vp],
    "dmc_batch_step": [_vp, C.c_int32],
    "dmc_batch_run": [_vp, C.c_int32],
    "dmc_batch_sync": [_vp, _i32],
    "dmc_batch_set_wpg": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "dmc_batch_tune": [_vp, C.c_int32, _i32, _f64],
}

_lib = None


def load():
    """Load libdmc_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise DMCError(2, f"{LIBPATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.dmc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(code):
    if code != 0:
        raise DMCError(code, load().dmc_last_error().decode(errors="replace"))
