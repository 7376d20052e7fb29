"""The device contract's host-side parts, after test/test_device_api.jl and
test/test_backend.jl, on the CPU: byte formatting, host-slab views, and the
ordering rules of queues and events on the jittered backend (a queue runs in
order, another queue sees its work only through an event, a synchronized
event or queue has finished, and failed queued work surfaces at the sync).
The same rules on the HIP backend are in test_gpu_backend_contract.py."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip.device_api import SLAB_ALIGN, human_readable_bytes, slab_matrix

from fakes import JitteryBackend, _NumpySlab, _np

C16 = np.dtype(np.complex128)


def test_human_readable_bytes():
    # test/test_device_api.jl:48-53
    assert human_readable_bytes(512) == "512 B"
    assert human_readable_bytes(1024) == "1.0 KiB"
    assert human_readable_bytes(3 * 2 ** 30) == "3.0 GiB"
    assert human_readable_bytes(5 * 2 ** 50).endswith("TiB")


def test_host_slab_views():
    # test/test_backend.jl:17-34 on a host stand-in slab
    slab = _NumpySlab(4096)
    assert slab.nbytes == 4096 and slab.ptr % SLAB_ALIGN == 0
    A = slab_matrix(slab, C16, (4, 8), 0)
    B = slab_matrix(slab, C16, (4, 8), 0)
    A[2, 4] = 1
    assert B[2, 4] == 1
    C = slab_matrix(slab, C16, (4, 8), 4 * 8 * 16)
    C.fill(0)
    assert A[2, 4] == 1
    with pytest.raises(F.ArgumentError):
        slab_matrix(slab, C16, (4, 8), 4000)
    with pytest.raises(F.ArgumentError):
        slab_matrix(slab, C16, (4, 8), -1)


def test_queue_ordering():
    # test/test_backend.jl:55-64: work queued on one queue runs in order
    b = JitteryBackend(jitter=2e-4)
    queue = b.make_queue()
    dev = b.alloc_device(np.float64, (1, 1))
    order = []

    def work():
        for i in range(64):
            b._launch(lambda i=i: order.append(i))
        b.fill(dev, 64.0)

    b.submit(queue, work)
    b.sync_queue(queue)
    assert order == list(range(64))
    assert _np(dev)[0, 0] == 64.0


def test_cross_queue_ordering_through_events():
    # test/test_backend.jl:66-80
    b = JitteryBackend(jitter=2e-4)
    for rep in range(10):
        up, down = b.make_queue(), b.make_queue()
        src = np.full((5, 3), rep + 1, dtype=C16, order="F")
        dev = b.alloc_device(C16, (5, 3))
        out = np.zeros((5, 3), dtype=C16, order="F")
        b.h2d(dev, src, queue=up)
        b.wait_event(down, b.record_event(up))
        b.d2h(out, dev, queue=down)
        b.sync_queue(down)
        assert np.array_equal(out, src)


def test_event_and_queue_synchronization():
    # test/test_backend.jl:82-98
    b = JitteryBackend(jitter=2e-4)
    queue = b.make_queue()
    counter = [0]

    def bump():
        counter[0] += 1

    b.submit(queue, lambda: [b._launch(bump) for _ in range(8)])
    b.sync_event(b.record_event(queue))
    assert counter[0] == 8
    b.submit(queue, lambda: b._launch(bump))
    b.sync_all()
    assert counter[0] == 9


def test_failed_queue_work_surfaces():
    # test/test_backend.jl:100-106
    b = JitteryBackend(jitter=2e-4)
    queue = b.make_queue()

    def boom():
        raise RuntimeError("boom")

    b.submit(queue, lambda: b._launch(boom))
    with pytest.raises(RuntimeError, match="boom"):
        b.sync_queue(queue)
