"""ctypes binding of libqedspin.so (include/qedh.h).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C qed_spin_amd/csrc``). There
is no fallback: a missing library or device raises, so nothing silently runs on the CPU.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

from . import errors

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqedspin.so")

_lock = threading.Lock()
_lib = None

c_int_p = C.POINTER(C.c_int)
c_i64_p = C.POINTER(C.c_int64)
c_u64_p = C.POINTER(C.c_uint64)
c_u32_p = C.POINTER(C.c_uint32)
c_i32_p = C.POINTER(C.c_int32)
c_dbl_p = C.POINTER(C.c_double)


class MatvecInfo(C.Structure):
    _fields_ = [
        ("rows", C.c_int64),
        ("cols", C.c_int64),
        ("is_real", C.c_int),
        ("lane", C.c_int),
        ("nnz", C.c_int64),
        ("value_bytes", C.c_int),
        ("id_width_ref", C.c_int),
        ("dict_size", C.c_int64),
        ("csr_bytes", C.c_uint64),
        ("build_seconds", C.c_double),
        ("lanes", C.c_int),
        ("norm_bound", C.c_double),
    ]


# name -> argument types (every entry point returns int status)
_SIGNATURES = {
    "qedh_device_count": [c_int_p],
    "qedh_device_memory": [c_u64_p, c_u64_p],
    "qedh_malloc": [C.POINTER(C.c_void_p), C.c_size_t],
    "qedh_free": [C.c_void_p],
    "qedh_memcpy_htod": [C.c_void_p, C.c_void_p, C.c_size_t],
    "qedh_memcpy_dtoh": [C.c_void_p, C.c_void_p, C.c_size_t],
    "qedh_memcpy_dtod": [C.c_void_p, C.c_void_p, C.c_size_t],
    "qedh_memset": [C.c_void_p, C.c_int, C.c_size_t],
    "qedh_synchronize": [],
    "qedh_thread_stream": [C.c_int],
    "qedh_vdot": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, c_dbl_p],
    "qedh_stream_stats": [c_u64_p, c_i64_p],
    "qedh_fill_zero": [C.c_int, C.c_void_p, C.c_size_t],
    "qedh_axpy": [C.c_int, c_dbl_p, C.c_void_p, C.c_void_p, C.c_size_t],
    "qedh_scale": [C.c_int, c_dbl_p, C.c_void_p, C.c_size_t],
    "qedh_axpby": [C.c_int, c_dbl_p, C.c_void_p, c_dbl_p, C.c_void_p, C.c_size_t],
    "qedh_dot": [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_nrm2": [C.c_int, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_axpy_dot": [C.c_int, c_dbl_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_axpy_nrm2": [C.c_int, c_dbl_p, C.c_void_p, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_dot_many": [C.c_int, C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_axpy_many": [C.c_int, c_dbl_p, C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p, C.c_size_t],
    "qedh_dot_many_noalloc": [C.c_int, C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p, C.c_size_t, c_dbl_p],
    "qedh_gemm": [
        C.c_int, C.c_char, C.c_char, C.c_size_t, C.c_size_t, C.c_size_t, c_dbl_p, C.c_void_p, C.c_size_t,
        C.c_void_p, C.c_size_t, c_dbl_p, C.c_void_p, C.c_size_t,
    ],
    "qedh_operator_create": [C.c_int, C.POINTER(C.c_void_p)],
    "qedh_operator_add_product": [C.c_void_p, C.c_char_p, c_int_p, C.c_double, C.c_double],
    "qedh_operator_info": [C.c_void_p, c_i64_p, c_dbl_p, c_int_p],
    "qedh_operator_destroy": [C.c_void_p],
    "qedh_program_compile": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "qedh_program_compile_projected": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
    "qedh_program_counts": [C.c_void_p, c_i64_p, c_i64_p, c_i64_p],
    "qedh_program_export": [C.c_void_p, c_u64_p, c_i32_p, c_u32_p, c_u64_p, c_u32_p, c_u64_p, c_dbl_p],
    "qedh_program_import": [
        C.c_int64, c_u64_p, c_i32_p, c_u32_p, C.c_int64, c_u64_p, c_u32_p, C.c_int64, c_u64_p, c_dbl_p,
        C.POINTER(C.c_void_p),
    ],
    "qedh_program_destroy": [C.c_void_p],
    "qedh_sector_import": [
        C.c_int, C.c_int, C.c_int, c_int_p, c_u64_p, c_dbl_p, C.c_int, C.c_int, C.c_int64, c_u64_p, c_dbl_p, C.c_uint64,
        C.POINTER(C.c_void_p),
    ],
    "qedh_sector_create": [
        C.c_int, C.c_int, C.c_int, c_int_p, c_u64_p, c_dbl_p, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)
    ],
    "qedh_sublattice_tables": [
        C.c_int, C.c_int, c_int_p, c_u64_p, C.c_int, C.c_int, c_int_p, c_int_p, C.c_void_p, c_u32_p, C.c_void_p,
        C.c_int64, c_i64_p,
    ],
    "qedh_sector_create_irrep": [
        C.c_int, C.c_int, C.c_int, c_int_p, c_u64_p, C.c_int, c_dbl_p, C.c_int, C.c_int, C.c_uint64,
        C.POINTER(C.c_void_p),
    ],
    "qedh_sector_irrep_info": [C.c_void_p, c_int_p, c_i64_p, c_i64_p],
    "qedh_sector_irrep_export": [C.c_void_p, C.c_void_p, c_u64_p, C.c_void_p, c_dbl_p],
    "qedh_matrix_elements": [
        C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p),
        C.POINTER(C.c_void_p), c_dbl_p,
    ],
    "qedh_sector_info": [C.c_void_p, c_i64_p, c_i64_p, c_int_p, c_dbl_p],
    "qedh_sector_export": [C.c_void_p, c_u64_p, c_dbl_p],
    "qedh_sector_key_bits": [C.c_void_p, c_int_p, c_int_p],
    "qedh_sector_index_and_projection": [C.c_void_p, c_u64_p, C.c_int64, c_i64_p, c_dbl_p],
    "qedh_sector_expand": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, c_dbl_p],
    "qedh_sector_destroy": [C.c_void_p],
    "qedh_matvec_create": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)],
    "qedh_matvec_info": [C.c_void_p, C.POINTER(MatvecInfo)],
    "qedh_matvec_apply": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int],
    "qedh_matvec_apply_host": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int],
    "qedh_matvec_apply_multi": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, C.c_int],
    "qedh_matvec_apply_block": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int],
    "qedh_matvec_set_lanes": [C.c_void_p, C.c_int],
    "qedh_matvec_set_tuning": [C.c_void_p, C.c_int, C.c_int],
    "qedh_matvec_export_csr": [C.c_void_p, c_u64_p, c_u32_p, c_dbl_p],
    "qedh_matvec_export_csr_rows": [C.c_void_p, C.c_int64, C.c_int64, c_u64_p, c_u32_p, c_dbl_p],
    "qedh_matvec_destroy": [C.c_void_p],
    "qedh_matvec_benchmark": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, c_dbl_p],
    "qedh_lanczos": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, c_dbl_p, c_dbl_p, c_int_p, c_dbl_p],
    "qedh_block_lowest": [C.c_void_p, C.c_int, C.c_int64, c_dbl_p, c_int_p, c_int_p, c_i64_p],
    "qedh_block_estimate": [C.c_void_p, c_dbl_p, c_dbl_p, C.POINTER(C.c_int64)],
    "qedh_block_eigenpairs": [C.c_void_p, C.c_int, C.c_int64, c_dbl_p, C.c_void_p, c_int_p, c_int_p],
    "qedh_ftlm": [
        C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_int, c_dbl_p, C.c_double, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p
    ],
    "qedh_cf": [
        C.c_void_p, c_dbl_p, C.c_int, C.c_int, c_dbl_p, C.c_double, C.c_int, C.c_double, C.c_double, c_dbl_p,
        c_dbl_p, c_dbl_p, c_dbl_p,
    ],
    "qedh_batch_lanczos": [
        C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, c_dbl_p, c_dbl_p, C.c_int, C.c_int, c_dbl_p, c_dbl_p,
        c_int_p, c_dbl_p,
    ],
    "qedh_ftlm_batched": [
        C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, c_u64_p, c_dbl_p, C.c_int, c_dbl_p, C.c_int, c_dbl_p,
        c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
    ],
    "qedh_cf_batched": [
        C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_double, C.c_double,
        C.c_int, c_dbl_p, C.c_double, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p, c_dbl_p,
    ],
    "qedh_serve": [C.c_void_p, C.c_int, C.c_void_p, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p, c_i64_p, c_dbl_p],
    "qedh_request_footprint": [C.c_void_p, C.c_int, c_u64_p],
    "qedh_ftlm_dynamics": [
        C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_uint64, C.c_double,
        C.c_int, c_dbl_p, C.c_int, c_dbl_p, C.c_double, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p,
    ],
    "qedh_ftlm_dynamics_walk": [
        C.c_int, C.POINTER(C.c_void_p), c_u64_p, c_dbl_p, C.c_int, c_int_p, C.POINTER(C.c_void_p),
        C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, c_dbl_p, C.c_int, c_dbl_p, C.c_double, C.c_uint64, c_dbl_p,
        c_dbl_p, c_dbl_p, c_int_p, c_dbl_p,
    ],
    "qedh_gaussian_vectors": [C.c_uint64, C.c_uint64, C.c_int, C.c_int64, C.c_void_p],
    "qedh_block_spectrum": [C.POINTER(C.c_void_p), C.c_int, C.c_int64, c_dbl_p],
    "qedh_exact_thermal": [
        C.POINTER(C.c_void_p), C.c_int, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
    ],
    "qedh_thermal_combine": [
        C.c_int, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
    ],
}

EXPORTED = sorted(list(_SIGNATURES) + ["qedh_last_error"])


def lib():
    """The loaded library (raises when it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise errors.DeviceUnavailable(
                    f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950) first"
                )
            L = C.CDLL(LIB_PATH)
            for name, args in _SIGNATURES.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = C.c_int
            L.qedh_last_error.argtypes = []
            L.qedh_last_error.restype = C.c_char_p
            _lib = L
        return _lib


def check(status: int) -> None:
    if status == 0:
        return
    msg = lib().qedh_last_error().decode(errors="replace")
    raise errors.from_status(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def device_count() -> int:
    n = C.c_int(0)
    call("qedh_device_count", C.byref(n))
    return int(n.value)


def device_memory():
    """(free, total) bytes of the current device."""
    f, t = C.c_uint64(0), C.c_uint64(0)
    call("qedh_device_memory", C.byref(f), C.byref(t))
    return int(f.value), int(t.value)


def ptr(a, ctype):
    """A ctypes pointer into a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
