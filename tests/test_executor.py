"""The executor's own tests, after test/test_executor.jl, on the CPU.

Every sweep runs on the jittered numpy backend (tests/fakes.py): its queues
are worker threads that sleep a random time before each operation, so only
the events the sweep waits on order the uploads, the panel function's work
and the write-backs, as on the device. A panel function reads its buffer with
backend.to_host, which waits for the work already queued on the sweep's
stream, as the reference's host(panel) does. Panels and steps are numbered
from 0.
"""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip.executor import PanelSteps, RowSteps, nsteps, readpanels, rowtraversal, stage_channel, sweep
from funicular_hip.executor import updatepanels, writepanels
from funicular_hip.plan import panel_bytes, row_block_height

from fakes import JitteryBackend

SWEEP_WIDTHS = (23, 8, 7, 1)  # test/test_executor.jl:1
C16, C8 = np.dtype(np.complex128), np.dtype(np.complex64)
REPETITIONS = 10


def randmatrix(rng, N, k):
    return np.asfortranarray(rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k)))


def make_plan(backend=None, **kw):
    backend = backend if backend is not None else JitteryBackend(jitter=0.0)
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20, **kw)


def jittery(seed=0):
    return JitteryBackend(jitter=2e-4, seed=seed)


def cols(pm, j):
    r = F.panelrange(pm, j)
    return slice(r[0], r[-1] + 1) if isinstance(r, range) else r


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_read_only_sweep_sees_every_panel(w, nb):
    # test/test_executor.jl:3-21
    rng = np.random.default_rng(1)
    N, k = 9, 23
    A = randmatrix(rng, N, k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
    order, seen, shapes = [], {}, {}

    def f(j, panel):
        order.append(j)
        seen[j] = backend.to_host(panel)
        shapes[j] = panel.shape

    F.foreachpanel(f, pm, write=False, nbuffers=nb)
    assert order == list(range(F.npanels(pm)))
    for j in order:
        assert shapes[j] == (N, F.panelwidth(pm, j))
        assert np.array_equal(seen[j], A[:, cols(pm, j)])
    assert np.array_equal(F.to_matrix(pm), A)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_read_write_sweep(w, nb):
    # test/test_executor.jl:23-36: panel .= 2 .* panel .+ j
    rng = np.random.default_rng(2)
    N, k = 9, 23
    A = randmatrix(rng, N, k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
    F.foreachpanel(lambda j, panel: backend.affine(panel, 2.0, float(j)), pm, nbuffers=nb)
    expected = 2 * A + np.array([c // w for c in range(k)], dtype=float)[None, :]
    assert np.array_equal(F.to_matrix(pm), expected)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_produce_sweep(w, nb):
    # test/test_executor.jl:38-52: target .= 3 .* source
    rng = np.random.default_rng(3)
    N, k = 9, 23
    A = randmatrix(rng, N, k)
    backend = jittery()
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)

    def f(j, source, target):
        backend.convert_copy(target, source)
        backend.scale(target, 3.0)

    sweep(f, readpanels(X), writepanels(Y), nbuffers=nb)
    assert np.array_equal(F.to_matrix(Y), 3 * A)
    assert np.array_equal(F.to_matrix(X), A)


def test_narrowed_host_storage_streams_through_the_sweep():
    # test/test_executor.jl:54-69
    rng = np.random.default_rng(4)
    N, k, w = 9, 20, 6
    A = randmatrix(rng, N, k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend, host_eltype=C8), w=w)
    narrowed = A.astype(C8).astype(C16)
    dtypes = []

    def f(j, panel):
        dtypes.append(panel.dtype)
        backend.scale(panel, 2.0)

    F.foreachpanel(f, pm)
    assert set(dtypes) == {C16}
    assert np.array_equal(F.to_matrix(pm), (2 * narrowed).astype(C8).astype(C16))


def test_narrowing_stages_through_buffers_out_of_the_pool():
    # test/test_executor.jl:71-93: only the narrow bytes cross the bus, so a
    # sweep over a narrowed matrix holds two buffers of each eltype and
    # allocates neither in its loop; the row traversal converts whole blocks.
    rng = np.random.default_rng(5)
    N, k, w = 9, 20, 6
    backend = JitteryBackend(jitter=0.0)
    plan = make_plan(backend, host_eltype=C8)
    pm = F.PanelMatrix.from_array(randmatrix(rng, N, k), plan=plan, w=w)
    held = 2 * (panel_bytes(N, w, C16) + panel_bytes(N, w, C8))
    F.foreachpanel(lambda j, panel: None, pm)
    assert F.device_bytes_allocated(plan) == held
    F.foreachpanel(lambda j, panel: None, pm)
    assert F.device_bytes_allocated(plan) == held
    before = F.to_matrix(pm)
    sweep(lambda b, block: backend.scale(block, 2.0), updatepanels(pm), traversal=rowtraversal(pm))
    assert np.array_equal(F.to_matrix(pm), (2 * before).astype(C8).astype(C16))


@pytest.mark.parametrize("nb", (1, 2))
def test_one_panels_buffer_never_leaks_into_another(nb):
    # test/test_executor.jl:95-113
    rng = np.random.default_rng(6)
    N, k, w = 6, 23, 5
    A = randmatrix(rng, N, k)
    stamped = np.array([[c // w for c in range(k)]] * N, dtype=C16)
    for rep in range(REPETITIONS):
        backend = jittery(rep)
        pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
        seen = {}

        def f(j, panel):
            seen[j] = backend.to_host(panel)
            backend.fill(panel, float(j))

        F.foreachpanel(f, pm, nbuffers=nb)
        assert all(np.array_equal(seen[j], A[:, cols(pm, j)]) for j in range(F.npanels(pm)))
        assert np.array_equal(F.to_matrix(pm), stamped)


def test_produce_sweep_keeps_its_input_and_output_buffers_apart():
    # test/test_executor.jl:115-131
    rng = np.random.default_rng(7)
    N, k, w = 6, 23, 5
    A = randmatrix(rng, N, k)
    for rep in range(REPETITIONS):
        backend = jittery(rep)
        plan = make_plan(backend)
        X = F.PanelMatrix.from_array(A, plan=plan, w=w)
        Y = F.PanelMatrix(N, k, plan=plan, w=w)
        sources = {}

        def f(j, source, target):
            backend.fill(target, float(j))
            sources[j] = backend.to_host(source)

        sweep(f, readpanels(X), writepanels(Y))
        assert all(np.array_equal(sources[j], A[:, cols(X, j)]) for j in range(F.npanels(X)))
        assert np.array_equal(F.to_matrix(X), A)


@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_nbuffers_1_and_2_move_data_bitwise_alike(w):
    # test/test_executor.jl:133-146
    rng = np.random.default_rng(8)
    N, k = 12, 23
    A = randmatrix(rng, N, k)
    results = []
    for nb in (1, 2):
        backend = jittery(nb)
        pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)

        def f(j, panel):
            backend.scale(panel, 0.3 + 0.7j)
            backend.affine(panel, 1.5, 0.25j)

        F.foreachpanel(f, pm, nbuffers=nb)
        results.append(F.to_matrix(pm))
    assert results[0].tobytes() == results[1].tobytes()


@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_nbuffers_1_and_2_reduce_bitwise_alike(w):
    # test/test_executor.jl:148-162
    rng = np.random.default_rng(9)
    N, k = 12, 23
    A = randmatrix(rng, N, k)
    sums = []
    for nb in (1, 2):
        backend = jittery(nb)
        pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
        total = [0j]

        def f(j, panel):
            total[0] += backend.to_host(panel).sum()

        F.foreachpanel(f, pm, write=False, nbuffers=nb)
        sums.append(total[0])
    assert sums[0] == sums[1]


def test_dirty_and_epoch_bookkeeping():
    # test/test_executor.jl:164-188
    N, k, w = 8, 20, 6
    backend = jittery()
    plan = make_plan(backend)
    pm = F.PanelMatrix(N, k, plan=plan, w=w)
    assert all(p.epoch == 0 and not p.dirty for p in pm.panels)
    F.foreachpanel(lambda j, panel: None, pm, write=False)
    assert all(p.epoch == 0 and not p.dirty for p in pm.panels)
    F.foreachpanel(lambda j, panel: backend.fill(panel, 0.0), pm)
    assert all(p.epoch == 1 and p.dirty for p in pm.panels)
    F.foreachpanel(lambda j, panel: None, pm)
    assert all(p.epoch == 2 for p in pm.panels)
    X = F.PanelMatrix(N, k, plan=plan, w=w)
    sweep(lambda j, source, target: backend.convert_copy(target, source), readpanels(X), writepanels(pm))
    assert all(p.epoch == 0 and not p.dirty for p in X.panels)
    assert all(p.epoch == 3 for p in pm.panels)


def test_sweeps_give_their_buffers_back_to_the_pool():
    # test/test_executor.jl:190-214
    rng = np.random.default_rng(10)
    N, k, w = 8, 20, 5
    backend = jittery()
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(randmatrix(rng, N, k), plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    assert F.device_bytes_allocated(plan) == 0
    pair = 2 * panel_bytes(N, w, C16)
    F.foreachpanel(lambda j, panel: None, X, write=False)
    assert F.device_bytes_allocated(plan) == pair
    F.foreachpanel(lambda j, panel: None, X)
    assert F.device_bytes_allocated(plan) == pair
    for _ in range(2):
        sweep(lambda j, source, target: backend.convert_copy(target, source), readpanels(X), writepanels(Y))
        assert F.device_bytes_allocated(plan) == 2 * pair
    assert np.array_equal(F.to_matrix(Y), F.to_matrix(X))


def test_nbuffers_is_capped_at_the_panel_count():
    # test/test_executor.jl:216-223
    rng = np.random.default_rng(11)
    plan = make_plan()
    pm = F.PanelMatrix.from_array(randmatrix(rng, 8, 20), plan=plan, w=20)
    assert F.npanels(pm) == 1
    F.foreachpanel(lambda j, panel: None, pm, nbuffers=2)
    assert F.device_bytes_allocated(plan) == panel_bytes(8, 20, C16)


def test_a_failing_panel_function_surfaces_and_frees_its_buffers():
    # test/test_executor.jl:225-238
    rng = np.random.default_rng(12)
    N, k, w = 8, 20, 5
    backend = jittery()
    plan = make_plan(backend)
    A = randmatrix(rng, N, k)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)

    def boom(j, panel):
        if j == 2:
            raise RuntimeError("boom")

    with pytest.raises(RuntimeError, match="boom"):
        F.foreachpanel(boom, pm)
    assert F.device_bytes_allocated(plan) == 2 * panel_bytes(N, w, C16)
    assert all(p.pins == 0 for p in pm.panels)
    F.foreachpanel(lambda j, panel: None, pm, write=False)
    assert F.device_bytes_allocated(plan) == 2 * panel_bytes(N, w, C16)
    assert np.array_equal(F.to_matrix(pm), A)


def test_sweep_validation():
    # test/test_executor.jl:240-256
    plan = make_plan()
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    Y = F.PanelMatrix(10, 20, plan=plan, w=7)
    Z = F.PanelMatrix(10, 20, plan=plan, w=5)
    noop = lambda j, *panels: None
    with pytest.raises(F.ArgumentError):
        sweep(noop)
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(Z))
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(X))
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), nbuffers=0)
    other = F.PanelMatrix(10, 20, plan=make_plan(), w=7)
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(other))
    assert sweep(noop, readpanels(X), writepanels(Y)) is None


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_a_panel_sweep_leaves_the_row_counts_free(w, nb):
    # test/test_executor.jl:258-274
    rng = np.random.default_rng(13)
    k = 23
    A = randmatrix(rng, 9, k)
    backend = jittery()
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(17, k, plan=plan, w=w)
    shapes = []

    def f(j, source, target):
        shapes.append((source.shape, (9, F.panelwidth(X, j)), target.shape, (17, F.panelwidth(Y, j))))
        backend.fill(target, float(j))

    sweep(f, readpanels(X), writepanels(Y), nbuffers=nb)
    assert all(a == b and c == d for a, b, c, d in shapes)
    assert np.array_equal(F.to_matrix(X), A)
    assert np.array_equal(F.to_matrix(Y), np.array([[c // w for c in range(k)]] * 17, dtype=C16))


def test_a_panel_sweep_refuses_matrices_cut_differently():
    # test/test_executor.jl:276-285
    plan = make_plan()
    X = F.PanelMatrix(9, 20, plan=plan, w=7)
    noop = lambda j, *panels: None
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(F.PanelMatrix(17, 20, plan=plan, w=5)))
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(F.PanelMatrix(17, 18, plan=plan, w=7)))
    assert sweep(noop, readpanels(X), writepanels(F.PanelMatrix(17, 20, plan=plan, w=7))) is None


@pytest.mark.parametrize("h", (1, 3, 5, 20))
@pytest.mark.parametrize("N", (12, 13, 1))
def test_row_blocks_tile_the_rows_exactly(N, h):
    # test/test_executor.jl:287-293
    t = RowSteps(N, h)
    assert [r for a, b in t.blocks for r in range(a, b)] == list(range(N))
    assert all(b - a <= min(h, N) for a, b in t.blocks)
    assert t.blocks[0][1] - t.blocks[0][0] == min(h, N)
    assert nsteps(t) == -(-N // min(h, N))


def test_row_block_height_matches_the_panel_buffer():
    # test/test_executor.jl:295-301
    assert row_block_height(120, 24, 24) == 120
    assert row_block_height(120, 24, 6) == 30
    assert row_block_height(120, 24, 1) == 5
    assert row_block_height(9, 23, 1) == 1
    assert row_block_height(120, 24, 100) == 120


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_read_only_row_sweep_sees_the_whole_matrix(w, nb):
    # test/test_executor.jl:303-319
    rng = np.random.default_rng(14)
    N, k = 13, 23
    A = randmatrix(rng, N, k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
    traversal = rowtraversal(pm)
    seen = np.zeros((N, k), dtype=C16)
    shapes = []

    def f(b, block):
        a, e = traversal.blocks[b]
        shapes.append((block.shape, (e - a, k)))
        seen[a:e, :] = backend.to_host(block)

    sweep(f, readpanels(pm), traversal=traversal, nbuffers=nb)
    assert all(s == t for s, t in shapes)
    assert np.array_equal(seen, A)
    assert np.array_equal(F.to_matrix(pm), A)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_read_write_row_sweep(w, nb):
    # test/test_executor.jl:321-336: rdiv_upper! on each row block
    rng = np.random.default_rng(15)
    N, k = 13, 23
    A = randmatrix(rng, N, k)
    R = np.triu(randmatrix(rng, k, k)) + 4 * np.eye(k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
    factor = backend.todevice(R)
    work = backend.alloc_device(np.uint8, (backend.trsm_workspace_bytes(k, C16), 1))
    sweep(lambda b, block: backend.rdiv_upper(block, factor, work=work), updatepanels(pm),
          traversal=rowtraversal(pm), nbuffers=nb)
    expected = np.linalg.solve(R.T, A.T).T
    assert np.linalg.norm(F.to_matrix(pm) - expected) <= 1e-12 * np.linalg.norm(expected)


@pytest.mark.parametrize("w", SWEEP_WIDTHS)
def test_row_sweeps_move_data_the_same_way_at_nbuffers_1_and_2(w):
    # test/test_executor.jl:338-351
    rng = np.random.default_rng(16)
    N, k = 13, 23
    A = randmatrix(rng, N, k)
    results = []
    for nb in (1, 2):
        backend = jittery(nb)
        pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
        sweep(lambda b, block: backend.scale(block, 2.0), updatepanels(pm), traversal=rowtraversal(pm), nbuffers=nb)
        results.append(F.to_matrix(pm))
    assert results[0].tobytes() == results[1].tobytes()
    assert np.array_equal(results[0], 2 * A)


@pytest.mark.parametrize("nb", (1, 2))
def test_a_row_sweep_leaves_the_column_counts_free(nb):
    # test/test_executor.jl:353-369: target .= view(source, :, 1:12)
    rng = np.random.default_rng(17)
    N = 24
    A = randmatrix(rng, N, 20)
    backend = jittery()
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=7)
    Y = F.PanelMatrix(N, 12, plan=plan, w=5)
    traversal = RowSteps(N, min(row_block_height(N, 20, 7), row_block_height(N, 12, 5)))
    shapes = []

    def f(b, source, target):
        a, e = traversal.blocks[b]
        shapes.append((source.shape, (e - a, 20), target.shape, (e - a, 12)))
        backend.convert_copy(target, source.view(cols=(0, 12)))

    sweep(f, readpanels(X), writepanels(Y), traversal=traversal, nbuffers=nb)
    assert all(a == b and c == d for a, b, c, d in shapes)
    assert np.array_equal(F.to_matrix(Y), A[:, :12])


def test_a_row_sweep_refuses_matrices_with_different_rows():
    # test/test_executor.jl:371-379
    plan = make_plan()
    X = F.PanelMatrix(24, 20, plan=plan, w=7)
    noop = lambda b, *blocks: None
    with pytest.raises(F.ArgumentError):
        sweep(noop, readpanels(X), writepanels(F.PanelMatrix(25, 20, plan=plan, w=7)), traversal=rowtraversal(X))
    assert sweep(noop, readpanels(X), writepanels(F.PanelMatrix(24, 9, plan=plan, w=3)),
                 traversal=rowtraversal(X)) is None


def test_a_repeated_panel_is_refused_when_the_sweep_writes():
    # test/test_executor.jl:381-390
    rng = np.random.default_rng(18)
    plan = make_plan()
    pm = F.PanelMatrix.from_array(randmatrix(rng, 8, 20), plan=plan, w=5)
    repeated = PanelSteps([0, 1, 0, 1], [0, 1, 0, 1])
    noop = lambda label, *panels: None
    with pytest.raises(F.ArgumentError):
        sweep(noop, updatepanels(pm), traversal=repeated)
    assert sweep(noop, readpanels(pm), traversal=repeated) is None


def test_a_custom_traversal_labels_its_steps():
    # test/test_executor.jl:392-406
    rng = np.random.default_rng(19)
    N, k, w = 8, 20, 5
    A = randmatrix(rng, N, k)
    backend = jittery()
    pm = F.PanelMatrix.from_array(A, plan=make_plan(backend), w=w)
    traversal = PanelSteps([1, 0, 1], ["second", "first", "again"])
    labels, seen = [], []

    def f(label, panel):
        labels.append(label)
        seen.append((label, backend.to_host(panel)))

    sweep(f, readpanels(pm), traversal=traversal)
    assert labels == ["second", "first", "again"]
    for label, got in seen:
        assert np.array_equal(got, A[:, cols(pm, 0 if label == "first" else 1)])


def test_the_panel_feed_is_bounded_and_in_order():
    # test/test_executor.jl:408-421
    plan = make_plan()
    pm = F.PanelMatrix(6, 20, plan=plan, w=5)
    feed = stage_channel(PanelSteps(F.npanels(pm)), (readpanels(pm),), depth=2)
    try:
        assert [feed.take() for _ in range(F.npanels(pm))] == list(range(F.npanels(pm)))
    finally:
        feed.close()
    blocked = stage_channel(PanelSteps(F.npanels(pm)), (readpanels(pm),), depth=1)
    try:
        import time
        time.sleep(0.05)
        assert blocked.available() <= 2
    finally:
        blocked.close()
