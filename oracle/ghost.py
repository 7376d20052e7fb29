"""ctypes wrapper of oracle/ghost.c.

TEST INFRASTRUCTURE (see oracle/__init__.py).
"""
import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.OUT):
            _build.build()
        L = ctypes.CDLL(_build.OUT)
        L.oracle_mix.restype = ctypes.c_uint64
        L.oracle_mix.argtypes = [ctypes.c_uint64]
        L.oracle_exp.restype = ctypes.c_double
        L.oracle_exp.argtypes = [ctypes.c_double]
        L.oracle_log.restype = ctypes.c_double
        L.oracle_log.argtypes = [ctypes.c_double]
        L.oracle_tables_checksum.restype = ctypes.c_uint64
        L.oracle_tables_checksum.argtypes = []
        L.oracle_zig_edges.restype = None
        L.oracle_zig_edges.argtypes = [ctypes.c_void_p]
        L.oracle_colkey.restype = ctypes.c_uint64
        L.oracle_colkey.argtypes = [ctypes.c_uint64, ctypes.c_int64]
        L.oracle_normal.restype = ctypes.c_double
        L.oracle_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for name in ("oracle_ghost_complex", "oracle_ghost_real"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                          ctypes.c_uint64]
        _lib = L
    return _lib


def ghost(dtype, row0, nrows, col0, ncols, seed):
    """Rows row0.. (0-based) of columns col0.. (1-based) of a ghost matrix."""
    dtype = np.dtype(dtype)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    if np.issubdtype(dtype, np.complexfloating):
        out = np.empty((nrows, ncols), dtype=np.complex128, order="F")
        lib().oracle_ghost_complex(out.ctypes.data, row0, nrows, col0, ncols, seed)
    else:
        out = np.empty((nrows, ncols), dtype=np.float64, order="F")
        lib().oracle_ghost_real(out.ctypes.data, row0, nrows, col0, ncols, seed)
    return out.astype(dtype)


def matrix(dtype, N, k, seed):
    """The whole N x k ghost matrix."""
    return ghost(dtype, 0, N, 1, k, seed)


def mix(z):
    return lib().oracle_mix(int(z) & 0xFFFFFFFFFFFFFFFF)


def exp(x):
    return lib().oracle_exp(float(x))


def log(x):
    return lib().oracle_log(float(x))


def tables_checksum():
    return lib().oracle_tables_checksum()


def zig_edges():
    out = np.empty(257, dtype=np.float64)
    lib().oracle_zig_edges(out.ctypes.data)
    return out
