"""The allocation-free solve, fhip_trsm_runn_noalloc, against fhip_trsm_runn
through the C ABI: the same bits on two calls in a row (the second smaller,
so it reuses the library's device scratch left by the first), no device
allocation, and the size limit refused before any device work."""
import ctypes

import numpy as np
import pytest
import scipy.linalg

from funicular_hip import _lib
from funicular_hip.device_api import eltype_code

pytestmark = pytest.mark.gpu


def _upper(rng, k, dtype):
    """A well-conditioned upper triangular factor (diagonal dominant)."""
    R = np.triu(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))) * 0.1
    R[np.diag_indices(k)] = 2.0 + rng.random(k)
    return np.asfortranarray(R.astype(dtype))


def _rows(rng, m, k, dtype):
    return np.asfortranarray((rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(dtype))


def _solve_both(backend, Y, R, compute):
    """Y R^-1 by fhip_trsm_runn (with a workspace) and by the allocation-free
    entry point, each on its own device copy of Y; returns both results and
    the counters read around the allocation-free call."""
    T = Y.dtype
    m, k = Y.shape
    Rd = backend.todevice(R)
    Y1, Y2 = backend.todevice(Y), backend.todevice(Y)
    need = backend.trsm_workspace_bytes(k, compute)
    work = backend.alloc_device(np.uint8, (need, 1))
    code, ccode = eltype_code(T), eltype_code(compute)
    _lib.call("fhip_trsm_runn", Y1.ptr, Y1.ld, m, k, Rd.ptr, Rd.ld, code, ccode, work.ptr, need, None)
    _lib.call("fhip_device_sync")
    before = _lib.counters()
    _lib.call("fhip_trsm_runn_noalloc", Y2.ptr, Y2.ld, m, k, Rd.ptr, Rd.ld, code, ccode, None)
    _lib.call("fhip_device_sync")
    after = _lib.counters()
    return backend.to_host(Y1), backend.to_host(Y2), before, after


@pytest.mark.parametrize("compute", [np.complex128, np.complex64])
def test_noalloc_solve_matches_the_workspace_solve(backend, compute):
    rng = np.random.default_rng(2024)
    T = np.complex128
    # The second call is smaller in both rows and columns: it runs on what the
    # first left in the shared scratch.
    for m, k in ((40000, 300), (5003, 70)):
        Y, R = _rows(rng, m, k, T), _upper(rng, k, T)
        ref, got, before, after = _solve_both(backend, Y, R, compute)
        assert ref.tobytes() == got.tobytes(), (m, k)
        assert after["device_allocs"] == before["device_allocs"]
        assert after["device_bytes"] == before["device_bytes"]
        # And it is the solve: Y R^-1 to the compute precision.
        X = scipy.linalg.solve_triangular(R, Y.T, trans="T", lower=False).T
        tol = 1e-12 if compute == np.complex128 else 1e-5
        assert np.linalg.norm(got - X) <= tol * np.linalg.norm(X)


def test_noalloc_solve_real_eltype_and_the_size_limit(backend):
    rng = np.random.default_rng(7)
    m, k = 7000, 129
    Y = np.asfortranarray(rng.standard_normal((m, k)))
    R = np.asfortranarray(np.triu(rng.standard_normal((k, k))) * 0.1 + np.diag(2.0 + rng.random(k)))
    ref, got, _, _ = _solve_both(backend, Y, R, np.float64)
    assert ref.tobytes() == got.tobytes()
    # k = 1024 in complex FP64 fills the 16 MiB scratch; k = 1025 does not fit
    # and is refused before anything touches the device.
    lib = _lib.load()
    assert lib.fhip_trsm_runn_noalloc(ctypes.c_void_p(256), 8, 8, 1025, ctypes.c_void_p(256), 1025,
                                      _lib.FHIP_C64, _lib.FHIP_C64, None) == _lib.FHIP_E_ARGUMENT
    assert b"scratch" in lib.fhip_last_error()
