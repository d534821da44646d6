"""ctypes binding of libspyker_hip.so (the C ABI declared in include/spyker_hip.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or cannot be loaded, every call raises.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "libspyker_hip.so")

I8, I16, I32, I64, U8, U16, U32, U64, F16, F32, F64 = range(11)

_i32, _i64, _f64, _p = C.c_int32, C.c_int64, C.c_double, C.c_void_p


class spk_winner(C.Structure):
    _fields_ = [("c", C.c_int64), ("t", C.c_int64), ("z", C.c_int64), ("y", C.c_int64), ("x", C.c_int64)]


class spk_stdp_config(C.Structure):
    _fields_ = [("positive", C.c_double), ("negative", C.c_double), ("stabilize", C.c_int32),
                ("lower", C.c_double), ("upper", C.c_double)]


class spk_sparse(C.Structure):
    _fields_ = [("n", C.c_int64), ("t", C.c_int64), ("c", C.c_int64), ("h", C.c_int64), ("w", C.c_int64),
                ("total", C.c_int64), ("offsets", C.c_void_p), ("events", C.c_void_p)]


# name -> argument types (all return int32 status unless listed in _RESTYPE)
SIGNATURES = {
    "spk_last_error": [],
    "spk_version": [],
    "spk_device_count": [_p],
    "spk_set_device": [_i32],
    "spk_current_device": [_p],
    "spk_memory_info": [_p, _p],
    "spk_synchronize": [_p],
    "spk_alloc": [_p, _i64],
    "spk_dealloc": [_p],
    "spk_memory_taken": [_p],
    "spk_memory_used": [_p],
    "spk_cache_clear": [],
    "spk_copy": [_p, _p, _i64, _p],
    "spk_threshold": [_p, _i32, _i64, _f64, _f64, _p],
    "spk_rank_fire": [_p, _i32, _p, _i32, _i64, _f64, _p],
    "spk_cast": [_p, _i32, _p, _i32, _i64, _p],
    "spk_fill": [_p, _i32, _i64, _f64, _p],
    "spk_pad": [_p, _p, _i32, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _f64, _p],
    "spk_rank_pool": [_p, _p, _i32, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _p],
    "spk_rank_inhibit": [_p, _i32, _i64, _i64, _i64, _i64, _f64, _p],
    "spk_rank_convwta": [_p, _i32, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _f64, _p, _p, _p],
    "spk_rank_fcwta": [_p, _i32, _i64, _i64, _i64, _i64, _i64, _f64, _p, _p, _p],
    "spk_conv": [_p, _i32, _p, _i32, _p, _i32] + [_i64] * 13 + [_p],
    "spk_fc": [_p, _i32, _p, _p, _i64, _i64, _i64, _p],
    "spk_dog": [_p, _p, _p] + [_i64] * 9 + [_p],
    "spk_gabor": [_p, _p, _p] + [_i64] * 9 + [_p],
    "spk_log": [_p, _p, _p] + [_i64] * 9 + [_p],
    "spk_rate_code": [_p, _i32, _p, _i32, _i64, _i64, _i64, _i32, _p],
    "spk_poisson_table": [_p, _i64, _i64],
    "spk_poisson_clear": [],
    "spk_rate_fire": [_p, _i32, _p, _i32, _i64, _i64, _i64, _f64, _p],
    "spk_zca_fit": [_p, _p, _p, _i64, _i64, _f64, _i32, _p],
    "spk_zca_trans": [_p, _p, _p, _i64, _i64, _p],
    "spk_zca_split": [_p, _p, _i32, _i64, _i64, _p],
    "spk_rank_code": [_p, _i32, _p, _i32, _i64, _i64, _i64, _i32, _p],
    "spk_rank_convstdp": [_p, _p, _i32, _p] + [_i64] * 10 + [_p, _i64, _p, _p, _i64] + [_i64] * 6 + [_p],
    "spk_rank_fcstdp": [_p, _p, _i32, _p, _i64, _i64, _i64, _i64, _p, _i64, _p, _p, _i64, _p],
    "spk_rank_gather": [_p, _i32, _p, _i32, _i64, _i64, _i64, _f64, _p],
    "spk_rank_scatter": [_p, _i32, _p, _i32, _i64, _i64, _i64, _p],
    "spk_rate_gather": [_p, _i32, _p, _i32, _i64, _i64, _i64, _p],
    "spk_sparse_release": [_p],
    "spk_sparse_from_dense": [_p, _i32, _i64, _i64, _i64, _i64, _i64, _f64, _p, _p],
    "spk_sparse_code": [_p, _i32, _i64, _i64, _i64, _i64, _i64, _i32, _p, _p],
    "spk_sparse_to_dense": [_p, _p, _p],
    "spk_sparse_gather": [_p, _p, _i32, _p],
    "spk_sparse_pad": [_p, _i64, _i64, _i64, _i64, _p, _p],
    "spk_sparse_pool": [_p] + [_i64] * 8 + [_p, _p],
    "spk_sparse_inhibit": [_p, _p, _p],
    "spk_sparse_convwta": [_p, _i64, _i64, _i64, _p, _p, _p],
    "spk_sparse_fcwta": [_p, _i64, _i64, _p, _p, _p],
    "spk_sparse_conv": [_p, _p, _i32, _i64, _i64, _i64, _f64] + [_i64] * 6 + [_p, _p],
    "spk_sparse_convstdp": [_p, _p, _i32, _i64, _i64, _i64, _p, _i64, _p, _p, _i64] + [_i64] * 4 + [_p],
    "spk_stdp_layer_fused": [_p] + [_i64] * 5 + [_p] + [_i64] * 7 + [_f64] + [_i64] * 3 + [_p, _i64, _p, _p]
                            + [_i64] * 4 + [_p, _p, _p, _p],
    "spk_stdp_layer_fused_noalloc": [_p] + [_i64] * 5 + [_p] + [_i64] * 7 + [_f64] + [_i64] * 3 + [_p, _i64, _p, _p]
                                    + [_i64] * 4 + [_p, _p, _p, _p],
    "spk_layer_plan_create": [_p, _i64, _p] + [_i64] * 8 + [_f64] + [_i64] * 7 + [_i32],
    "spk_layer_plan_layout": [_p, _p, _p, _p, _p],
    "spk_layer_plan_run": [_p, _p, _p, _i32, _p, _i64, _p, _p, _p, _p, _p],
    "spk_layer_plan_candidates": [_p, _p, _p, _i32, _p, _p, _p],
    "spk_layer_plan_select": [_p, _i32, _i32],
    "spk_layer_plan_destroy": [_p],
    "spk_set_fused_arith": [_i32],
    "spk_set_fused_tuning": [_i32],
    "spk_set_fused_channels": [_i32],
    "spk_fused_tuned_channels": [_i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _p],
    "spk_kernel_timer": [_i32],
    "spk_kernel_timer_read": [_p, _p],
    "spk_random_seed": [_i64],
    "spk_normal_kernel": [_p, _i64, _f64, _f64],
    "spk_gaussian_kernel": [_p, _i64, _f64],
    "spk_gabor_kernel": [_p, _i64, _f64, _f64, _f64, _f64, _f64],
}
_RESTYPE = {"spk_last_error": C.c_char_p}

_lib = None


def load():
    """Returns the loaded library, raising if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(PATH):
        raise RuntimeError("spyker_amd: %s is not built (python -m spyker_amd.build)" % PATH)
    lib = C.CDLL(PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int32)
    _lib = lib
    return lib


# spk_status values (include/spyker_hip.h)
ERR_ARGUMENT, ERR_UNSUPPORTED, ERR_RUNTIME, ERR_CONFIG = -1, -2, -3, -4


class SpykerError(RuntimeError):
    """A nonzero status from the library: the message and the spk_status code."""

    def __init__(self, msg, code):
        super().__init__(msg)
        self.code = code


def call(name, *args):
    """Calls one entry point; a nonzero status raises SpykerError (a RuntimeError) with the library's message."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.spk_last_error().decode(errors="replace")
        raise SpykerError("%s failed (%d): %s" % (name, rc, msg), rc)
    return rc
