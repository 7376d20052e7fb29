"""ctypes binding of libfunicular_hip.so (declared in include/funicular_hip.h).

The library is the only compute path: loading fails loudly when it is missing
or was not built, and every call that returns a nonzero status raises with the
library's own message. There is no fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libfunicular_hip.so")

# Status codes (include/funicular_hip.h)
FHIP_OK = 0
FHIP_E_ARGUMENT = 1
FHIP_E_DIMENSION = 2
FHIP_E_HIP = 3
FHIP_E_WORKSPACE = 4
FHIP_E_UNSUPPORTED = 5

# Element types
FHIP_F32 = 0
FHIP_F64 = 1
FHIP_C32 = 2
FHIP_C64 = 3

# Copy kinds
FHIP_H2D = 1
FHIP_D2H = 2
FHIP_D2D = 3
FHIP_H2H = 4


class Counters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "device_allocs", "device_bytes", "host_allocs", "host_bytes", "h2d_bytes",
        "d2h_bytes", "d2d_bytes", "copies", "kernel_launches", "host_syncs")]


class LibraryError(RuntimeError):
    """A call into libfunicular_hip.so failed; `code` is the FHIP_E_* status."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64

# name -> argument types; every function returns int status unless listed in _RESTYPES.
_SIGNATURES = {
    "fhip_last_error": [],
    "fhip_version": [],
    "fhip_device_count": [ctypes.POINTER(_i)],
    "fhip_set_device": [_i],
    "fhip_device_arch": [ctypes.c_char_p, _sz],
    "fhip_device_sync": [],
    "fhip_malloc": [ctypes.POINTER(_vp), _sz],
    "fhip_free": [_vp],
    "fhip_host_alloc": [ctypes.POINTER(_vp), _sz],
    "fhip_host_free": [_vp],
    "fhip_host_is_pinned": [_vp, _vp],
    "fhip_stream_create": [ctypes.POINTER(_vp)],
    "fhip_stream_destroy": [_vp],
    "fhip_stream_sync": [_vp],
    "fhip_event_create": [ctypes.POINTER(_vp), _i],
    "fhip_event_destroy": [_vp],
    "fhip_event_record": [_vp, _vp],
    "fhip_stream_wait_event": [_vp, _vp],
    "fhip_event_sync": [_vp],
    "fhip_event_query": [_vp, ctypes.POINTER(_i)],
    "fhip_event_elapsed": [ctypes.POINTER(ctypes.c_float), _vp, _vp],
    "fhip_memcpy2d_async": [_vp, _sz, _vp, _sz, _sz, _sz, _i, _vp],
    "fhip_get_counters": [ctypes.POINTER(Counters)],
    "fhip_reset_counters": [],
    "fhip_gram_workspace_bytes": [ctypes.POINTER(_sz), _i64, _i64, _i64, _i, _i, _i],
    "fhip_gram": [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _d, _d, _d, _d, _i, _i, _vp, _sz, _vp],
    "fhip_trsm_workspace_bytes": [ctypes.POINTER(_sz), _i64, _i],
    "fhip_trsm_runn": [_vp, _i64, _i64, _i64, _vp, _i64, _i, _i, _vp, _sz, _vp],
    "fhip_trsm_runc": [_vp, _i64, _i64, _i64, _vp, _i64, _i, _vp, _sz, _vp],
    "fhip_trsm_runn_noalloc": [_vp, _i64, _i64, _i64, _vp, _i64, _i, _i, _vp],
    "fhip_gemm_nn": [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _i, _vp],
    "fhip_convert_copy": [_vp, _i, _i64, _vp, _i, _i64, _i64, _i64, _vp],
    "fhip_fill": [_vp, _i, _i64, _i64, _i64, _d, _d, _vp],
    "fhip_rowdot": [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _i, _i, _vp],
    "fhip_axpby": [_vp, _i64, _vp, _i64, _i64, _i64, _d, _d, _d, _d, _i, _vp],
    "fhip_scale": [_vp, _i64, _i64, _i64, _d, _d, _i, _i, _vp],
    "fhip_affine": [_vp, _i64, _i64, _i64, _d, _d, _i, _d, _d, _i, _vp],
    "fhip_trsm_apply": [_vp, _i64, _i64, _i64, _i, _i, _i, _vp, ctypes.c_size_t, _vp],
    "fhip_potrf_upper": [_vp, _i64, _i, _vp, _i64, _i, _i64, _vp, _vp],
    "fhip_gemm_nn_scaled": [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _d, _d, _d, _d, _i, _vp],
    "fhip_dot_workspace_bytes": [ctypes.POINTER(ctypes.c_size_t), _i64, _i64],
    "fhip_dot": [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _i, _i, _i, _vp, ctypes.c_size_t, _vp],
    "fhip_ghost_randn": [_vp, _i, _i64, _i64, _i64, _i64, _i64, _u64, _vp],
    "fhip_ghost_tables_checksum": [ctypes.POINTER(_u64)],
    "fhip_stencil_tridiag": [_vp, _i64, _vp, _i64, _i64, _i64, _d, _d, _d, _d, _vp, _i, _i, _vp],
    "fhip_stencil_helmholtz3d": [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _d, _d, _d, _d, _i, _i, _vp],
}

_RESTYPES = {"fhip_last_error": ctypes.c_char_p, "fhip_version": ctypes.c_char_p}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load():
    """Loads the library once and binds every entry point of the header."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: build it with `python funicular.jl_amd/build.py` "
            "(or __graft_entry__.build()). The HIP library is the only compute path.")
    lib = ctypes.CDLL(LIBRARY_PATH)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ctypes.c_int)
    _lib = lib
    return lib


def check(status):
    """Raises LibraryError carrying fhip_last_error() for a nonzero status."""
    if status != FHIP_OK:
        lib = load()
        msg = lib.fhip_last_error().decode(errors="replace")
        raise LibraryError(status, msg or f"libfunicular_hip status {status}")
    return status


def call(name, *args):
    return check(getattr(load(), name)(*args))


def counters():
    c = Counters()
    call("fhip_get_counters", ctypes.byref(c))
    return {name: getattr(c, name) for name, _ in Counters._fields_}
