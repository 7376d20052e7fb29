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
