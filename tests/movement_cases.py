"""Panel movement cases after test/test_movement.jl:1-106, run on any
backend: whole-panel materialize and writeback with and without queues,
their refusals, narrowed host storage, and queued movement fenced by events.
Imported by the CPU file (the jittered fake backend) and the GPU file (the
HIP backend)."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip.panels import materialize, writeback

C16 = np.complex128


def randmatrix(seed, N, k, T=C16):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k))
    return np.asfortranarray(A.astype(T))


def plan_for(backend, **kw):
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20, **kw)


def synchronous_movement(backend, w):
    N, k = 11, 20
    A = randmatrix(1, N, k)
    plan = plan_for(backend)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    buffer = backend.alloc_device(C16, (N, w))
    for j in range(F.npanels(pm)):
        wj = F.panelwidth(pm, j)
        view = buffer.view(None, (0, wj))
        materialize(view, pm.panels[j], plan)
        assert np.array_equal(backend.to_host(view), A[:, F.panelrange(pm, j)])
        backend.scale(view, 2.0)
        writeback(pm.panels[j], view, plan)
    assert np.array_equal(F.to_matrix(pm), 2 * A)


def movement_rejects_mismatched_buffers(backend):
    plan = plan_for(backend)
    pm = F.PanelMatrix(11, 20, plan=plan, w=7)
    last = pm.panels[-1]  # 6 columns
    with pytest.raises(F.DimensionMismatch):
        materialize(backend.alloc_device(C16, (11, 7)), last, plan)
    with pytest.raises(F.DimensionMismatch):
        writeback(last, backend.alloc_device(C16, (11, 7)), plan)
    assert backend.sync_event(materialize(backend.alloc_device(C16, (11, 6)), last, plan)) is None


def movement_with_narrowed_storage(backend):
    N, k = 9, 15
    A = randmatrix(2, N, k)
    plan = plan_for(backend, host_eltype=np.complex64)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=4)
    buffer = backend.alloc_device(C16, (N, 4))
    narrowed = A.astype(np.complex64).astype(C16)
    for j in range(F.npanels(pm)):
        view = buffer.view(None, (0, F.panelwidth(pm, j)))
        materialize(view, pm.panels[j], plan)
        assert np.array_equal(backend.to_host(view), narrowed[:, F.panelrange(pm, j)])
        writeback(pm.panels[j], view, plan)
    assert np.array_equal(F.to_matrix(pm), narrowed)


def queued_movement(backend):
    N, k, w = 12, 23, 5
    A = randmatrix(3, N, k)
    plan = plan_for(backend)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    up, compute, down = backend.make_queue(), backend.make_queue(), backend.make_queue()
    from funicular_hip import plan as P
    buffers = P.checkout_device_buffers(plan, C16, (N, w), 2)
    try:
        for j in range(F.npanels(pm)):
            slot = buffers[j % 2].view(None, (0, F.panelwidth(pm, j)))
            backend.wait_event(up, backend.record_event(down))
            materialize(slot, pm.panels[j], plan, queue=up)
            backend.wait_event(compute, backend.record_event(up))
            backend.submit(compute, lambda slot=slot: backend.scale(slot, 2.0))
            backend.wait_event(down, backend.record_event(compute))
            writeback(pm.panels[j], slot, plan, queue=down)
        for q in (up, compute, down):
            backend.sync_queue(q)
    finally:
        P.checkin_device_buffers(plan, buffers)
    assert np.array_equal(F.to_matrix(pm), 2 * A)
    assert all(p.epoch == 2 and p.dirty for p in pm.panels)


def queued_matches_synchronous(backend):
    N, k, w = 10, 17, 6
    A = randmatrix(4, N, k)
    plan = plan_for(backend)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    buffer = backend.alloc_device(C16, (N, w))
    reference = np.empty_like(A)
    for j in range(F.npanels(pm)):
        slot = buffer.view(None, (0, F.panelwidth(pm, j)))
        materialize(slot, pm.panels[j], plan)
        reference[:, F.panelrange(pm, j)] = backend.to_host(slot)
    queued = F.PanelMatrix.from_array(A, plan=plan, w=w)
    queued_buffer = backend.alloc_device(C16, (N, w))
    out = np.empty_like(A)
    queue = backend.make_queue()
    for j in range(F.npanels(queued)):
        slot = queued_buffer.view(None, (0, F.panelwidth(queued, j)))
        materialize(slot, queued.panels[j], plan, queue=queue)
        # A copy inside submit without a queue of its own lands on the
        # submitting queue, after the materialize queued there.

        def read(j=j, slot=slot):
            out[:, F.panelrange(queued, j)] = backend.to_host(slot)
        backend.submit(queue, read)
    backend.sync_queue(queue)
    assert np.array_equal(out, reference)
