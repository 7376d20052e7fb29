"""Plan arithmetic: the reference's own known answers (test/test_plan.jl,
test/test_executor.jl:287-301) for the oracle and for the backend's plan, and
the reference's properties (test/test_plan.jl:40-66) over seeded cases."""
import numpy as np
import pytest

from oracle import plan as OP
from funicular_hip import plan as P
from funicular_hip.device_api import ArgumentError
from funicular_hip.executor import RowSteps, nsteps
from fakes import HostOnlyBackend

C16 = np.complex128


def make_plan(**kw):
    kw.setdefault("device_budget", 64 * 2 ** 20)
    kw.setdefault("host_budget", 64 * 2 ** 20)
    return P.ResidencyPlan(backend=HostOnlyBackend(), **kw)


@pytest.mark.parametrize("impl", ["oracle", "backend"])
def test_panel_width_known_answers(impl):
    # test/test_plan.jl:28-47
    if impl == "oracle":
        cpw = lambda N, k, **kw: OP.choose_panel_width(N, k, 16, **kw)
        err = ValueError
    else:
        cpw = lambda N, k, **kw: P.choose_panel_width(N, k, C16, **kw)
        err = ArgumentError
    assert cpw(100, 64, device_budget=1 << 30) == 64
    assert cpw(100, 64, device_budget=2 * 2 * 100 * 16) == 2
    assert cpw(100, 1000, device_budget=2 * 999 * 100 * 16) == 500
    assert cpw(100, 64, device_budget=1 << 30, target_panel_bytes=8 * 100 * 16) == 8
    assert cpw(100, 64, device_budget=1 << 30, target_panel_bytes=1) == 1
    with pytest.raises(err):
        cpw(100, 64, device_budget=100 * 16)
    with pytest.raises(err):
        cpw(0, 64, device_budget=1 << 30)
    with pytest.raises(err):
        cpw(100, 0, device_budget=1 << 30)


def test_panel_bytes_known_answers():
    # test/test_plan.jl:19-21
    assert P.panel_bytes(1000, 8, C16) == 1000 * 8 * 16
    assert P.device_bytes_required(1000, 8, C16, nbuffers=2, workspace_bytes=7) == 2 * 1000 * 8 * 16 + 7
    assert OP.device_bytes_required(1000, 8, 16, nbuffers=2, workspace_bytes=7) == 2 * 1000 * 8 * 16 + 7


def test_panel_width_properties_and_agreement():
    # test/test_plan.jl:40-66 over this build's own seeded cases, plus the
    # backend's plan agreeing with the oracle on every one.
    rng = np.random.default_rng(0xDEADBEEF)
    sizes = {np.complex64: 8, np.complex128: 16, np.float64: 8}
    for _ in range(400):
        N = int(rng.integers(1, 5001))
        k = int(rng.integers(1, 2001))
        T = [np.complex64, np.complex128, np.float64][rng.integers(0, 3)]
        nb = int(rng.integers(1, 3))
        extra = int(rng.integers(0, 3))
        ws = int(rng.integers(0, (1 << 20) + 1))
        column = N * sizes[T]
        budget = ws + (nb + extra) * column * int(rng.integers(1, 2001))
        target = [1, 1 << 10, 1 << 20, 2 * 1024 ** 3][rng.integers(0, 4)]
        w = P.choose_panel_width(N, k, T, device_budget=budget, nbuffers=nb, extra_buffers=extra,
                                 workspace_bytes=ws, target_panel_bytes=target)
        wo = OP.choose_panel_width(N, k, sizes[T], budget, nb, extra, ws, target)
        assert w == wo
        per_buffer = (budget - ws) // (nb + extra)
        uncapped = min(k, per_buffer // column)
        capped = min(k, max(1, min(target, per_buffer) // column))
        assert 1 <= w <= uncapped
        assert P.device_bytes_required(N, w, T, nbuffers=nb, extra_buffers=extra, workspace_bytes=ws) <= budget
        assert OP.cld(k, w) == OP.cld(k, capped)
        assert w == 1 or P.panel_bytes(N, w, T) <= max(target, column)


@pytest.mark.parametrize("impl", ["oracle", "backend"])
def test_row_block_height_known_answers(impl):
    # test/test_executor.jl:295-301
    rbh = OP.row_block_height if impl == "oracle" else P.row_block_height
    assert rbh(120, 24, 24) == 120
    assert rbh(120, 24, 6) == 30
    assert rbh(120, 24, 1) == 5
    assert rbh(9, 23, 1) == 1
    assert rbh(120, 24, 100) == 120


@pytest.mark.parametrize("N", [12, 13, 1])
@pytest.mark.parametrize("h", [1, 3, 5, 20])
def test_row_blocks_tile_rows(N, h):
    # test/test_executor.jl:287-293
    for blocks in (OP.row_blocks(N, h), RowSteps(N, h).blocks):
        flat = [r for (a, b) in blocks for r in range(a, b)]
        assert flat == list(range(N))
        assert all(b - a <= min(h, N) for (a, b) in blocks)
        assert blocks[0][1] - blocks[0][0] == min(h, N)
        assert len(blocks) == OP.cld(N, min(h, N))
    assert nsteps(RowSteps(N, h)) == len(OP.row_blocks(N, h))


def test_configuration_schedules():
    # The row blocks and panel counts of the five configurations.
    cases = [(1048576, 64, 8, 8, 131072, 8), (10077696, 512, 32, 16, 629856, 16),
             (8388608, 256, 16, 16, 524288, 16), (4194304, 128, 32, 4, 1048576, 4),
             (33386248, 1024, 32, 32, 1043320, 33)]
    for N, k, w, p, h, nblocks in cases:
        assert len(OP.panel_ranges(k, w)) == p
        assert OP.row_block_height(N, k, w) == h == P.row_block_height(N, k, w)
        assert len(OP.row_blocks(N, h)) == nblocks
    last = OP.row_blocks(33386248, 1043320)[-1]
    assert last[1] - last[0] == 8


def test_plan_construction():
    # test/test_plan.jl:1-16
    plan = make_plan(scratch_dir="/tmp/funicular-test", workspace_bytes=2 ** 20)
    assert plan.device_budget == 64 * 2 ** 20
    assert plan.devicepool.budget == 64 * 2 ** 20 - 2 ** 20
    assert plan.scratch_dir == "/tmp/funicular-test"
    assert "ResidencyPlan" in repr(plan)
    for kw in (dict(nbuffers=0), dict(host_budget=0), dict(workspace_bytes=-1),
               dict(device_budget=2 ** 20, workspace_bytes=2 ** 20), dict(panel_width=0),
               dict(target_panel_bytes=0)):
        with pytest.raises(ArgumentError):
            make_plan(**kw)


def test_plan_width_resolution():
    # test/test_plan.jl:68-88
    plan = make_plan(device_budget=2 * 64 * 100 * 16)
    assert P.resolve_panel_width(plan, 100, 256, C16) == 64
    assert P.resolve_panel_width(plan, 100, 256, C16, override=17) == 17
    assert P.resolve_panel_width(plan, 100, 256, C16, extra_buffers=2) == 32
    for o in (0, 257, 128):
        with pytest.raises(ArgumentError):
            P.resolve_panel_width(plan, 100, 256, C16, override=o)
    fixed = make_plan(device_budget=1 << 30, panel_width=13)
    assert P.resolve_panel_width(fixed, 100, 256, C16) == 13
    assert P.resolve_panel_width(fixed, 100, 256, C16, override=8) == 8
    with pytest.raises(ArgumentError) as e:
        P.resolve_panel_width(plan, 100, 256, C16, override=128)
    assert "device_budget" in str(e.value)
    assert "panel_width ≤ 64" in str(e.value)


def test_device_pool():
    # test/test_plan.jl:90-110
    plan = make_plan(device_budget=5 * 64 * 16 + 64)
    bufs = P.checkout_device_buffers(plan, C16, (64, 1), 2)
    assert len(bufs) == 2 and bufs[0] is not bufs[1]
    assert P.device_bytes_allocated(plan) == 2 * 64 * 16
    P.checkin_device_buffers(plan, bufs)
    again = P.checkout_device_buffers(plan, C16, (64, 1), 2)
    assert {id(b) for b in again} == {id(b) for b in bufs}
    assert P.device_bytes_allocated(plan) == 2 * 64 * 16
    nested = P.checkout_device_buffers(plan, C16, (64, 1), 3)
    assert not ({id(b) for b in nested} & {id(b) for b in again})
    assert P.device_bytes_allocated(plan) == 5 * 64 * 16
    with pytest.raises(ArgumentError):
        P.checkout_device_buffers(plan, C16, (64, 1), 1)
    narrow = P.checkout_device_buffers(plan, np.complex64, (4, 1), 1)
    assert narrow[0].dtype == np.complex64


def test_host_pool():
    # test/test_plan.jl:112-134
    plan = make_plan(host_budget=4 * 2 ** 20)
    pool = plan.hostpool
    blocks = [P.alloc_host_storage(plan, C16, (2 ** 16, 1)) for _ in range(4)]
    assert all(s.array.shape == (2 ** 16, 1) for s in blocks)
    assert P.host_bytes_reserved(pool) == pool.budget
    assert P.host_bytes_in_use(pool) == 4 * 2 ** 20
    assert P.host_bytes_free(pool) == 0
    with pytest.raises(ArgumentError):
        P.alloc_host_storage(plan, C16, (2 ** 16, 1))
    P.release_host_block(plan, blocks[1].block)
    assert P.host_bytes_free(pool) == 2 ** 20
    assert P.host_bytes_in_use(pool) == 3 * 2 ** 20
    reused = P.alloc_host_storage(plan, C16, (2 ** 16, 1))
    assert reused.block == blocks[1].block
    assert P.host_bytes_free(pool) == 0
    blocks[2].array[:] = 1
    blocks[3].array[:] = 2
    assert blocks[2].block.slab == blocks[3].block.slab
    assert np.all(blocks[2].array == 1)


def test_host_pool_grows_in_slabs():
    # test/test_plan.jl:136-147
    plan = make_plan(host_budget=8 * 2 ** 20)
    pool = plan.hostpool
    assert P.host_bytes_reserved(pool) == 0
    P.alloc_host_storage(plan, C16, (2 ** 16, 1))
    assert P.host_bytes_reserved(pool) == 2 ** 20
    for _ in range(7):
        P.alloc_host_storage(plan, C16, (2 ** 16, 1))
    assert P.host_bytes_reserved(pool) == 8 * 2 ** 20
    assert len(pool.slabs) == 4


def test_project_schedule():
    # src/panelblas.jl:272-281: p (p + 1) steps, labels in order.
    from funicular_hip.panelblas import project_schedule
    for p in (1, 2, 4, 16):
        panels, labels = project_schedule(p)
        op, ol = OP.project_labels(p)
        assert panels == op and labels == ol
        assert len(panels) == p * (p + 1)
    panels, labels = project_schedule(2)
    assert labels == [(True, 0, 0), (False, 0, 0), (False, 1, 0), (True, 1, 1), (False, 0, 1), (False, 1, 1)]
