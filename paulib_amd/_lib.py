"""ctypes binding of libpaulib_hip.so (the C ABI in include/paulib_hip.h).

The library is loaded from paulib_amd/_build/; if it is missing or fails to
load, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "libpaulib_hip.so"

PLH_OK = 0
ERRORS = {
    -1: "PLH_ERR_ARG",
    -2: "PLH_ERR_AS",
    -3: "PLH_ERR_SIZE",
    -4: "PLH_ERR_HIP",
    -5: "PLH_ERR_OOM",
    -6: "PLH_ERR_NOGPU",
    -7: "PLH_ERR_STATE",
}


class EngineError(RuntimeError):
    """A non-zero status from the C ABI; .code is the PLH_ERR_* value."""

    def __init__(self, code: int, func: str, msg: str):
        super().__init__(f"{func}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


_p = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_i64 = C.c_int64
_int = C.c_int

# name -> (restype, argtypes); mirrors include/paulib_hip.h line by line.
SIGNATURES = {
    "plh_last_error": (C.c_char_p, []),
    "plh_version": (_int, []),
    "plh_build_info": (C.c_char_p, []),
    "plh_ctx_create": (_int, [_int, _p, C.POINTER(_p)]),
    "plh_ctx_destroy": (_int, [_p]),
    "plh_ctx_synchronize": (_int, [_p]),
    "plh_ctx_stream": (_p, [_p]),
    "plh_ctx_scratch_bytes": (_i64, [_p]),
    "plh_ctx_autotune": (_int, [_p, C.c_char_p, _int]),
    "plh_ctx_tuning": (_int, [_p, _i64p, _i64p, _i64p]),
    "plh_ctx_trim": (_int, [_p]),
    "plh_ctx_profile": (_int, [_p, _int]),
    "plh_ctx_profile_read": (_int, [_p, _int, C.c_char_p, _f64p, _i64p, _f64p, C.POINTER(_int)]),
    "plh_sum_create": (_int, [_p, _int, _i64, _i64, C.POINTER(_p)]),
    "plh_sum_destroy": (_int, [_p]),
    "plh_sum_upload": (_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p]),
    "plh_sum_download": (_int, [_p, _u64p, _u64p, _u8p, _f64p]),
    "plh_sum_download_range": (_int, [_p, _i64, _i64, _u64p, _u64p, _u8p, _f64p]),
    "plh_sum_length": (_i64, [_p]),
    "plh_sum_num_qubits": (_i64, [_p]),
    "plh_sum_set_num_qubits": (_int, [_p, _i64]),
    "plh_sum_array_size": (_int, [_p]),
    "plh_sum_device_planes": (_int, [_p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), _i64p]),
    "plh_sum_copy": (_int, [_p, _p]),
    "plh_launch_mul_soa": (_int, [_int, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _int, _int, _p]),
    "plh_multiply": (_int, [_p, _p]),
    "plh_multiply_canonical": (_int, [_p, _p, _p]),
    "plh_batch_multiply_canonical": (_int, [_int, C.POINTER(_p), C.POINTER(_p), _p, _i64p]),
    "plh_batch_multiply_canonical_each": (_int, [_int, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), _i64]),
    "plh_commutator": (_int, [_p, _p, _int, _p]),
    "plh_sub": (_int, [_p, _p]),
    "plh_scale": (_int, [_p, C.c_double]),
    "plh_lookup": (_int, [_p, _int, _p, _p, C.c_uint8, C.POINTER(C.c_double)]),
    "plh_trace": (_int, [_p, C.POINTER(C.c_double)]),
    "plh_frobenius_norm2": (_int, [_p, C.POINTER(C.c_double)]),
    "plh_trim_keep": (_int, [_p, _i64, _p]),
    "plh_apply_truncation_keep": (_int, [_p, _i64, C.c_double, _i64, _p]),
    "plh_launch_commutator_soa_noalloc": (_int, [_int, _int, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                                 _i64, _i64, _i64, _p]),
    "plh_commutator_canonical": (_int, [_p, _p, _int, _p]),
    "plh_sort_and_combine": (_int, [_p]),
    "plh_canonicalize": (_int, [_p]),
    "plh_add": (_int, [_p, _p]),
    "plh_zero_state_expectation": (_int, [_p, _f64p]),
    "plh_truncate_weight": (_int, [_p, _i64]),
    "plh_cutoff": (_int, [_p, C.c_double]),
    "plh_apply_H": (_int, [_p, _i64]),
    "plh_apply_S": (_int, [_p, _i64]),
    "plh_apply_CNOT": (_int, [_p, _i64, _i64]),
    "plh_apply_clifford": (_int, [_p, C.c_int, _i64, _i64]),
    **{f"plh_apply_{g}": (_int, [_p, _i64]) for g in ("I", "X", "Y", "Z", "H_YZ", "H_XY", "sqrt_X", "sqrt_Y")},
    **{f"plh_apply_{g}": (_int, [_p, _i64, _i64])
       for g in ("CZ", "CY", "XC_X", "XC_Z", "XC_Y", "YC_X", "YC_Z", "YC_Y", "SWAP", "iSWAP")},
    "plh_clifford_table": (_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "plh_apply_RX": (_int, [_p, _i64, C.c_double]),
    "plh_apply_RY": (_int, [_p, _i64, C.c_double]),
    "plh_apply_RZ": (_int, [_p, _i64, C.c_double]),
    "plh_elementwise_multiply": (_int, [_p, _p]),
    "plh_evolve": (_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p]),
    "plh_apply_truncation": (_int, [_p, _i64, C.c_double, _i64]),
    "plh_trim": (_int, [_p, _i64]),
    "plh_propagate": (_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p, _i64, C.c_double, _i64]),
    "plh_run_program": (_int, [_p, _i64, _i32p, _i64p, _i64p, _u64p, _u64p, _f64p, _i64, C.c_double, _i64]),
    "plh_server_create": (_int, [_p, _int, _int, _i64, _int, C.POINTER(_p)]),
    "plh_server_destroy": (_int, [_p]),
    "plh_server_submit": (_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p, _i64, _i64, _i32p, _i64p, _i64p,
                                 _u64p, _u64p, _f64p, _i64, C.c_double, _i64, _i64p]),
    "plh_server_step": (_int, [_p, C.POINTER(_int), C.POINTER(_int)]),
    "plh_server_poll": (_int, [_p, _int, _i64p, _f64p, _i64p, _f64p, C.POINTER(_int)]),
    "plh_server_take_terms": (_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p]),
    "plh_server_batch_terms": (_i64, [_p]),
    "plh_server_steps": (_i64, [_p]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded engine; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP engine first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, func: str) -> None:
    if rc != PLH_OK:
        msg = lib().plh_last_error()
        raise EngineError(rc, func, msg.decode() if msg else "")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc
