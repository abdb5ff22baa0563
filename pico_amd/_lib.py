"""ctypes binding of libpico_gpu.so (the C ABI in include/pico_gpu.h).

The library is loaded from this package directory.  There is no fallback: if
the shared library is missing or fails to load, importing the package raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpico_gpu.so")

PG_OK = 0
PG_E_INVALID = -1
PG_E_OOM = -2
PG_E_DEVICE = -3
PG_E_ADMISSION = -4
PG_E_UNSUPPORTED = -5
PG_E_CANCELLED = -6

KOALABEAR = 0
BABYBEAR = 1


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pico_gpu error {code}: {msg}")
        self.code = code


class AdmissionError(PgError):
    pass


class Challenger(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int32), ("state", ctypes.c_uint32 * 16),
                ("input", ctypes.c_uint32 * 8), ("n_input", ctypes.c_int32),
                ("output", ctypes.c_uint32 * 8), ("n_output", ctypes.c_int32)]


_P = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_i = ctypes.c_int
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every symbol of include/pico_gpu.h
SIGNATURES = {
    "pg_last_error": (ctypes.c_char_p, []),
    "pg_version": (_i, []),
    "pg_create": (_i, [_i, _sz, _pp]),
    "pg_destroy": (None, [_P]),
    "pg_bytes_in_use": (_sz, [_P]),
    "pg_device_synchronize": (_i, [_i]),
    "pg_profile_enable": (_i, [_P, _i]),
    "pg_profile_reset": (_i, [_P]),
    "pg_profile_get": (_i, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_uint64)]),
    "pg_debug_fail_alloc": (_i, [_P, ctypes.c_longlong]),
    "pg_debug_host_blocks": (_i, [_P, ctypes.c_size_t, ctypes.c_size_t]),
    "pg_autotune": (_i, [_P, _i, _i, _i, _P, _sz, _u32, _i, ctypes.c_char_p, _sz]),
    "pg_tune_set": (_i, [_i, _i, _i, _i, _i]),
    "pg_tune_get": (_i, [_i, _i, _i, _u32, ctypes.POINTER(ctypes.c_int)]),
    "pg_tune_export": (_i, [ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_size_t)]),
    "pg_tune_import": (_i, [ctypes.c_char_p]),
    "pg_buf_alloc": (_i, [_P, _sz, _sz, _pp]),
    "pg_buf_wrap": (_i, [_P, _P, _sz, _sz, _pp]),
    "pg_buf_free": (None, [_P]),
    "pg_host_alloc": (_i, [_P, _sz, _pp]),
    "pg_host_free": (None, [_P]),
    "pg_buf_device_ptr": (_P, [_P]),
    "pg_buf_rows": (_sz, [_P]),
    "pg_buf_cols": (_sz, [_P]),
    "pg_buf_ld": (_sz, [_P]),
    "pg_buf_view": (_i, [_P, _P, _sz, _sz, _sz, _sz, _sz, _pp]),
    "pg_upload": (_i, [_P, _P, _P, _i, _i]),
    "pg_download": (_i, [_P, _P, _sz, _sz, _P, _i, _i]),
    "pg_fill_random": (_i, [_P, _P, _i, _u64]),
    "pg_coset_lde": (_i, [_P, _i, _P, _u32, _u32, _i, _pp]),
    "pg_mmcs_commit": (_i, [_P, _i, _P, _sz, _i, _P, _pp]),
    "pg_mmcs_open_batch": (_i, [_P, _P, _P, _sz, _P, _P]),
    "pg_mmcs_verify_batch": (_i, [_i, _P, _P, _sz, _P, _u64, _P, _P]),
    "pg_tree_root": (_i, [_P, _P]),
    "pg_tree_num_matrices": (_sz, [_P]),
    "pg_tree_matrix": (_P, [_P, _sz]),
    "pg_tree_log_height": (_i, [_P]),
    "pg_tree_total_width": (_sz, [_P]),
    "pg_tree_digest_layer": (_i, [_P, _P, _i, _P]),
    "pg_tree_free": (None, [_P]),
    "pg_pcs_commit": (_i, [_P, _i, _P, _P, _sz, _u32, _P, _pp]),
    "pg_pcs_commit_host": (_i, [_P, _i, _P, _P, _P, _P, _sz, _u32, _i, _P, _pp]),
    "pg_pcs_evaluations": (_i, [_P, _P, _sz, _sz, _P]),
    "pg_pcs_open_values": (_i, [_P, _P, _sz, _u32, _P, _sz, _P]),
    "pg_pcs_open_values_noalloc": (_i, [_P, _P, _sz, _u32, _P, _sz, _P]),
    "pg_pcs_reduce_openings": (_i, [_P, _P, _sz, _P, _P, _P, _P, _P]),
    "pg_pcs_open": (_i, [_P, _P, _sz, _P, _P, _u32, _i, _sz, _i, _P, _pp]),
    "pg_open_proof_sizes": (_i, [_P, _P, _P, _P, _P, _P, _P]),
    "pg_open_proof_read": (_i, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pg_open_proof_encode": (_i, [_P, _i, _P, _sz, _P]),
    "pg_open_proof_decode": (_i, [_i, _P, _sz, _i, _pp]),
    "pg_open_proof_free": (None, [_P]),
    "pg_air_compile": (_i, [_i, _P, _sz, _P, _sz, _pp, _P, _P]),
    "pg_air_program_free": (None, [_P]),
    "pg_quotient_values": (_i, [_P, _P, _P, _sz, _P, _sz, _u32, _u32, _P, _sz, _P, _pp]),
    "pg_quotient_values_perm": (_i, [_P, _P, _P, _sz, _P, _sz, _P, _sz, _u32, _u32, _P, _sz, _P, _sz, _P, _P, _sz,
                                     _P, _pp]),
    "pg_batch_create": (_i, [_P, _i, _P, _P, _P, _sz, _i, _pp]),
    "pg_batch_run": (_i, [_P, _P, _P]),
    "pg_batch_take_tree": (_i, [_P, _sz, _pp]),
    "pg_batch_num_launches": (_sz, [_P]),
    "pg_batch_free": (None, [_P]),
    "pg_service_create": (_i, [_P, _sz, _i, _pp]),
    "pg_service_submit": (_i, [_P, _i, _P, _u32, _P]),
    "pg_service_submit_synthetic": (_i, [_P, _i, _u64, _sz, _sz, _u32, _P]),
    "pg_service_submit_host": (_i, [_P, _i, _P, _sz, _sz, _i, _u32, _P]),
    "pg_service_wait": (_i, [_P, _u64, _P, _P]),
    "pg_service_open": (_i, [_P, _u64, _P, _sz, _P, _P]),
    "pg_service_release": (_i, [_P, _u64]),
    "pg_service_cancel": (_i, [_P, _u64]),
    "pg_service_stats": (_i, [_P, _P, _P, _P]),
    "pg_service_destroy": (None, [_P]),
    "pg_fri_fold": (_i, [_P, _i, _P, _P, _pp]),
    "pg_fri_commit_phase": (_i, [_P, _i, _P, _P, _i, _i, _i, _P, _P, _P, _P, _pp]),
    "pg_fri_grind": (_i, [_P, _i, _P, _i, _P]),
    "pg_fri_answer_queries": (_i, [_P, _P, _P, _sz, _P, _P]),
    "pg_fri_num_rounds": (_i, [_P]),
    "pg_fri_round_tree": (_P, [_P, _i]),
    "pg_fri_free": (None, [_P]),
    "pg_challenger_init": (None, [_P, _i]),
    "pg_challenger_observe": (None, [_P, _u32]),
    "pg_challenger_sample": (_u32, [_P]),
    "pg_challenger_sample_ext": (None, [_P, _P]),
    "pg_challenger_sample_bits": (_u32, [_P, _i]),
    "pg_challenger_check_witness": (_i, [_P, _i, _u32]),
}

_lib = None


def lib():
    """Load the HIP library (once).  Raises if it is missing: no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C pico_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != PG_OK:
        msg = lib().pg_last_error().decode(errors="replace")
        if rc == PG_E_ADMISSION:
            raise AdmissionError(rc, msg)
        raise PgError(rc, msg)
    return rc
