"""ctypes binding of ``libtnmps.so`` (the C ABI declared in include/tnmps.h).

The shared library is built in-tree next to this file by
``tensor-net_amd/build.py``.  There is no fallback: if the library is missing
or no HIP device is visible, creating a state raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtnmps.so"

# tn_status values (include/tnmps.h)
TN_OK = 0
TN_EMPTY_SYSTEM = 1
TN_INVALID_QUANTUM_NUMBER = 2
TN_UNPHYSICAL_INDEX = 3
TN_STATE_INCOMPATIBLE_SHAPE = 4
TN_OPERATOR_INCOMPATIBLE_SHAPE = 5
TN_NONE = 6
TN_ZERO_PROB = 7
TN_SAMPLE_RANGE = 8
TN_CAPACITY = 9
TN_DEVICE = 10
TN_INVALID_ARG = 11
TN_NO_CONVERGENCE = 12

# sticky failure flag bits (tn_take_flags)
TN_FLAG_NO_CONVERGENCE = 1
TN_FLAG_CAPACITY = 2

TRUNC_NONE, TRUNC_CONST, TRUNC_CUTOFF = 0, 1, 2
PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2


class tn_trunc(C.Structure):
    _fields_ = [("kind", C.c_int), ("chi", C.c_size_t), ("eps", C.c_double)]


class tn_tune(C.Structure):
    _fields_ = [("c_max", C.c_int), ("theta_tile", C.c_int), ("bw", C.c_int),
                ("split_max", C.c_int), ("split_rows", C.c_int), ("ahead", C.c_int),
                ("qr", C.c_int)]


TUNE_FIELDS = ("c_max", "theta_tile", "bw", "split_max", "split_rows", "ahead", "qr")
# fields a table may leave out, and the value they then take
TUNE_DEFAULTS = {"qr": 1}


class tn_perf(C.Structure):
    _fields_ = [("ms_theta", C.c_double), ("ms_jacobi", C.c_double),
                ("ms_post", C.c_double), ("ms_measure", C.c_double),
                ("alg_flops", C.c_double), ("svd_flops", C.c_double),
                ("updates", C.c_size_t), ("sweeps", C.c_size_t),
                ("jacobi_launches", C.c_size_t), ("ms_host_launch", C.c_double),
                ("factored", C.c_size_t), ("ms_split", C.c_double), ("ms_renorm", C.c_double),
                ("bytes_split", C.c_double), ("bytes_renorm", C.c_double),
                ("bytes_measure", C.c_double), ("measure_sites", C.c_size_t),
                ("ms_contract", C.c_double), ("contract_flops", C.c_double),
                ("qr_tasks", C.c_size_t), ("ms_qr", C.c_double)]


_sz = C.c_size_t
_szp = C.POINTER(C.c_size_t)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_h = C.c_void_p
_hp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every entry point of include/tnmps.h
SIGNATURES = {
    "tn_status_str": (C.c_char_p, [C.c_int]),
    "tn_last_error": (C.c_char_p, []),
    "tn_device_count": (C.c_int, []),
    "tn_mps_new_qubits": (C.c_int, [_sz, tn_trunc, _sz, C.c_void_p, _hp]),
    "tn_mps_free": (None, [_h]),
    "tn_mps_from_vector": (C.c_int, [_sz, _dp, tn_trunc, _sz, C.c_void_p, _hp]),
    "tn_engine_new": (C.c_int, [_sz, C.c_void_p, _hp]),
    "tn_engine_free": (None, [_h]),
    "tn_mps_new_qubits_in": (C.c_int, [_h, _sz, tn_trunc, _sz, _hp]),
    "tn_set_precision": (C.c_int, [_h, C.c_int]),
    "tn_engine_set_tuning": (C.c_int, [_h, C.POINTER(tn_tune), _sz]),
    "tn_engine_get_tuning": (C.c_int, [_h, C.POINTER(tn_tune), _sz, _szp]),
    "tn_set_tuning": (C.c_int, [_h, C.POINTER(tn_tune), _sz]),
    "tn_batch_bilayer": (C.c_int, [_h, _sz, _hp, _szp, _szp, _dp, _szp, _szp, _ip, _dp, _szp,
                                   _dp, _szp, _ip, _dp]),
    "tn_mps_clone": (C.c_int, [_h, _hp]),
    "tn_mps_n": (_sz, [_h]),
    "tn_get_trunc": (C.c_int, [_h, C.POINTER(tn_trunc)]),
    "tn_set_trunc": (C.c_int, [_h, tn_trunc]),
    "tn_sync": (C.c_int, [_h]),
    "tn_apply_u1": (C.c_int, [_h, _sz, _dp]),
    "tn_apply_u2": (C.c_int, [_h, _sz, _dp]),
    "tn_apply_u2_multi": (C.c_int, [_h, _sz, _sz, _dp]),
    "tn_apply_u2_noalloc": (C.c_int, [_h, _sz, _dp]),
    "tn_take_flags": (C.c_int, [_h, _ip]),
    "tn_apply_unis": (C.c_int, [_h, _sz, _szp, _ip, _dp]),
    "tn_prob": (C.c_int, [_h, _sz, _sz, _dp]),
    "tn_probs": (C.c_int, [_h, _sz, _dp]),
    "tn_measure_prob": (C.c_int, [_h, _sz, C.c_double, _ip, _dp]),
    "tn_measure_postsel_prob": (C.c_int, [_h, _sz, _sz, _dp]),
    "tn_meas_layer": (C.c_int, [_h, _sz, _szp, _ip, _dp, _szp, _szp, _ip, _dp]),
    "tn_apply_bilayer": (C.c_int, [_h, _sz, _szp, _ip, _dp, _sz, _szp, _ip, _dp, _szp, _szp,
                                   _ip, _dp]),
    "tn_entropy_vn": (C.c_int, [_h, _sz, _dp]),
    "tn_entropy_ry_schmidt": (C.c_int, [_h, C.c_double, _sz, _dp]),
    "tn_schmidt": (C.c_int, [_h, _sz, _dp, _sz, _szp]),
    "tn_bond_dims": (C.c_int, [_h, _szp]),
    "tn_max_bond_dim": (C.c_int, [_h, _szp]),
    "tn_get_gamma": (C.c_int, [_h, _sz, _dp, _sz, _szp]),
    "tn_run_circuit": (C.c_int, [_h, _sz, _szp, _szp, _dp, _szp, _szp, _ip, _dp, _sz, _dp,
                                 _szp, _ip, _dp]),
    "tn_stats": (C.c_int, [_h, _szp, _szp]),
    "tn_set_timing": (C.c_int, [_h, C.c_int]),
    "tn_get_perf": (C.c_int, [_h, C.POINTER(tn_perf)]),
    "tn_reset_perf": (C.c_int, [_h]),
    "tn_engine_stats": (C.c_int, [_h, _szp, _szp]),
    "tn_engine_set_timing": (C.c_int, [_h, C.c_int]),
    "tn_engine_get_perf": (C.c_int, [_h, C.POINTER(tn_perf)]),
    "tn_engine_reset_perf": (C.c_int, [_h]),
    "tn_engine_sync": (C.c_int, [_h]),
    "tn_entropies": (C.c_int, [_h, _sz, _hp, _szp, _dp, _dp]),
    "tn_contract": (C.c_int, [_h, _dp, _sz]),
    "tn_matrix_part": (C.c_int, [_h, _sz, _sz, _dp, _sz]),
    "tn_rng_words": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _sz,
                               C.POINTER(C.c_uint64)]),
    "tn_gen_haar4": (C.c_int, [_h, C.c_uint64, _sz, C.POINTER(C.c_uint64), _dp]),
    "tn_batch_bilayer_gen": (C.c_int, [_h, _sz, _hp, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64), _dp, _szp, _dp, _szp, _szp, _ip,
                                       _dp]),
}

_LIB = None


def load() -> C.CDLL:
    """Load the native library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.environ.get("TNMPS_LIB", str(LIB_PATH))
    if not Path(path).exists():
        raise ImportError(
            f"libtnmps.so not found at {path}: build it with "
            "`python tensor-net_amd/build.py` (hipcc, gfx950)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def status_str(code: int) -> str:
    return load().tn_status_str(code).decode()


def last_error() -> str:
    return load().tn_last_error().decode()
