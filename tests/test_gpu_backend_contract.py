"""What only a device backend can be held to, after test/test_cuda.jl:6-61:
its traits and queues, the page-locking of the host tier's slabs, and the
stream a panel function runs on."""
import ctypes

import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import _lib
from funicular_hip.device_api import Event, current_stream
from funicular_hip.panels import panelstorage

pytestmark = pytest.mark.gpu


def plan(backend):
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20)


def pinned(address):
    out = ctypes.c_int(-1)
    _lib.call("fhip_host_is_pinned", ctypes.byref(out), address)
    return out.value


def test_backend_traits(backend):
    assert backend.supports_eltype(np.complex128)
    assert backend.supports_eltype(np.float32)
    assert not backend.supports_eltype(np.int64)
    A = backend.alloc_device(np.complex128, (7, 3))
    assert A.dtype == np.dtype(np.complex128) and tuple(A.shape) == (7, 3)
    assert isinstance(backend.record_event(backend.make_queue()), Event)
    # A sweep gets three queues of its own, each a different stream, none of
    # them the null stream.
    up, compute, down = backend.make_sweep_queues()
    streams = {up.stream, compute.stream, down.stream}
    assert len(streams) == 3 and None not in streams and 0 not in streams


def test_host_slabs_are_page_locked(backend):
    pm = F.PanelMatrix(64, 8, plan=plan(backend), w=4)
    block = panelstorage(pm.panels[0])
    assert pinned(block.ctypes.data) == 1
    assert pinned(np.empty(1 << 16).ctypes.data) == 0


def test_panel_functions_run_on_the_sweeps_stream(backend):
    A = np.asfortranarray(np.arange(32 * 12, dtype=np.complex128).reshape(32, 12))
    pm = F.PanelMatrix.from_array(A, plan=plan(backend), w=5)
    outer = current_stream()
    seen = []
    F.foreachpanel(lambda j, panel: seen.append(current_stream()), pm, write=False)
    assert len(seen) == F.npanels(pm) == 3
    # One stream for the whole sweep, and not the caller's: a kernel a panel
    # function launches lands on the sweep's compute queue.
    assert all(s == seen[0] for s in seen)
    assert seen[0] is not None and seen[0] != outer
    assert current_stream() == outer


# The vendor contract of test/test_backend.jl on the HIP backend.

def test_host_slab(backend):
    # test/test_backend.jl:17-34: page-locked, aligned, views alias, bounds checked
    from funicular_hip.device_api import SLAB_ALIGN, slab_matrix
    slab = backend.alloc_host_slab(4096)
    assert slab.nbytes == 4096 and slab.ptr % SLAB_ALIGN == 0
    assert pinned(slab.ptr) == 1
    A = slab_matrix(slab, np.complex128, (4, 8), 0)
    B = slab_matrix(slab, np.complex128, (4, 8), 0)
    A[2, 4] = 1
    assert B[2, 4] == 1
    C = slab_matrix(slab, np.complex128, (4, 8), 4 * 8 * 16)
    C.fill(0)
    assert A[2, 4] == 1
    with pytest.raises(F.ArgumentError):
        slab_matrix(slab, np.complex128, (4, 8), 4000)
    with pytest.raises(F.ArgumentError):
        slab_matrix(slab, np.complex128, (4, 8), -1)


@pytest.mark.parametrize("T", (np.complex128, np.float64, np.complex64, np.float32), ids=str)
def test_synchronous_copies(backend, T):
    # test/test_backend.jl:36-53, narrowing to the next narrower eltype
    T = np.dtype(T)
    rng = np.random.default_rng(20)
    src = rng.standard_normal((6, 4))
    if T.kind == "c":
        src = src + 1j * rng.standard_normal((6, 4))
    src = np.asfortranarray(src.astype(T))
    dev = backend.alloc_device(T, (6, 4))
    assert backend.h2d(dev, src) is None
    assert np.array_equal(backend.to_host(dev), src)
    out = np.zeros((6, 4), dtype=T, order="F")
    backend.d2h(out, dev)
    assert np.array_equal(out, src)
    S = {np.dtype(np.complex128): np.complex64, np.dtype(np.float64): np.float32}.get(T)
    if S is not None:
        narrow = backend.alloc_device(S, (6, 4))
        backend.h2d(narrow, src)
        backend.d2h(out, narrow)
        assert np.array_equal(out, src.astype(S).astype(T))
    with pytest.raises(F.DimensionMismatch):
        backend.h2d(backend.alloc_device(T, (6, 5)), src)


def test_cross_queue_ordering_through_events(backend):
    # test/test_backend.jl:66-80: the download queue sees the upload only
    # through the event it waits on
    for rep in range(10):
        up, down = backend.make_queue(), backend.make_queue()
        src = np.full((1 << 16, 3), rep + 1, dtype=np.complex128, order="F")
        dev = backend.alloc_device(np.complex128, src.shape)
        out = np.zeros_like(src)
        backend.h2d(dev, src, queue=up)
        backend.wait_event(down, backend.record_event(up))
        backend.d2h(out, dev, queue=down)
        backend.sync_queue(down)
        assert np.array_equal(out, src)


def test_event_and_queue_synchronization(backend):
    # test/test_backend.jl:82-98: kernels queued inside submit run on the
    # queue; a synchronized event, or sync_all, has seen them finish
    queue = backend.make_queue()
    dev = backend.alloc_device(np.float64, (1 << 20, 1))
    backend.fill(dev, 0.0)
    backend.sync_all()
    backend.submit(queue, lambda: [backend.affine(dev, 1.0, 1.0) for _ in range(8)])
    backend.sync_event(backend.record_event(queue))
    assert np.all(backend.to_host(dev) == 8.0)
    backend.submit(queue, lambda: backend.affine(dev, 1.0, 1.0))
    backend.sync_all()
    assert np.all(backend.to_host(dev) == 9.0)


def test_failed_queue_work_surfaces(backend):
    # test/test_backend.jl:100-106: work submitted to a queue runs on the
    # calling thread, so its failure surfaces at the submit
    queue = backend.make_queue()

    def boom():
        raise RuntimeError("boom")

    with pytest.raises(RuntimeError, match="boom"):
        backend.submit(queue, boom)
    # The queue is still usable afterwards.
    dev = backend.alloc_device(np.float64, (4, 1))
    backend.submit(queue, lambda: backend.fill(dev, 2.0))
    backend.sync_queue(queue)
    assert np.all(backend.to_host(dev) == 2.0)
