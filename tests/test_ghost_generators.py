"""GhostPanels with a host column generator, the reference's form
GhostPanels(f!, T, N, k) (src/ghosts.jl:53-122), on the CPU through the
jittered backend: the generator sees 1-based columns, the values do not
depend on the panel width or the visiting order, the seed changes them, and
a sweep reads them as an ordinary host-resident matrix
(test/test_ghosts.jl:1-56, 88-100 are the model)."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import panels as P
from fakes import JitteryBackend


def _plan(backend=None, **kw):
    backend = JitteryBackend(jitter=0.0) if backend is None else backend
    kw.setdefault("host_budget", 64 << 20)
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, **kw)


def test_the_generator_sees_one_based_columns():
    seen = []

    def f(dst, rng, col):
        seen.append(col)
        dst[...] = col
    plan = _plan()
    O = F.GhostPanels(f, np.float64, 7, 10, plan=plan, w=4)
    A = F.to_matrix(O)
    assert np.array_equal(A, np.tile(np.arange(1, 11, dtype=float), (7, 1)))
    assert sorted(set(seen)) == list(range(1, 11))


@pytest.mark.parametrize("T", [np.complex128, np.float64, np.complex64])
def test_width_independent_and_seed_sensitive(T):
    plan = _plan()
    ref = F.to_matrix(F.GhostPanels(F.randn_column, T, 300, 13, plan=plan, seed=0xDEADBEEF, w=13))
    for w in (1, 4, 7):
        got = F.to_matrix(F.GhostPanels(T, 300, 13, plan=plan, seed=0xDEADBEEF, w=w, generator=F.randn_column))
        assert np.array_equal(got, ref)
    other = F.to_matrix(F.GhostPanels(F.randn_column, T, 300, 13, plan=plan, seed=0xDEADBEEF + 1, w=4))
    assert not np.array_equal(other, ref)
    # Standard normal entries, complex parts of variance 1/2 each.
    v = np.var(ref.real) + (np.var(ref.imag) if np.iscomplexobj(ref) else 0.0)
    assert 0.8 < v < 1.2


def test_sweeps_read_host_generated_panels_in_any_order():
    # Row sweeps stage every panel; panel sweeps one at a time, regenerating
    # whatever the host budget let go.
    backend = JitteryBackend(jitter=1e-4, seed=3)
    plan = _plan(backend)
    O = F.GhostPanels(F.randn_column, np.complex128, 200, 12, plan=plan, seed=11, w=4)
    A = F.to_matrix(O)
    assert np.allclose(F.gram(O), A.conj().T @ A, rtol=1e-13, atol=1e-12)
    Y = F.PanelMatrix(200, 12, plan=plan, w=4)
    F.panelmul(Y, F.DenseOperator(np.eye(200), backend), O)
    assert np.array_equal(F.to_matrix(Y), A)


def test_a_ghost_with_a_host_generator_refuses_writes_and_bad_arguments():
    plan = _plan()
    O = F.GhostPanels(F.randn_column, np.complex128, 50, 6, plan=plan, w=3)
    with pytest.raises(F.ArgumentError):
        F.scale(O, 2.0)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(F.randn_column, np.complex128, 50, 6, plan=plan, generator=F.randn_column)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(np.complex128, 50, 6, plan=plan, generator=3)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(np.complex128, 0, 6, plan=plan)


def test_the_default_generator_is_randn_column_and_the_device_stream_is_named():
    # GhostPanels(T, N, k) takes the host generator, as the reference's takes
    # randn! (src/ghosts.jl:53-63); the device stream is chosen by name.
    plan = _plan()
    default = F.to_matrix(F.GhostPanels(np.complex128, 120, 9, plan=plan, seed=5, w=4))
    named = F.to_matrix(F.GhostPanels(F.randn_column, np.complex128, 120, 9, plan=plan, seed=5, w=4))
    assert np.array_equal(default, named)
    device = F.GhostPanels(F.device_randn, np.complex128, 120, 9, plan=plan, seed=5, w=4)
    assert all(p.home.source.generator is None for p in device.panels)
    assert repr(F.device_randn) == "device_randn"


# The rest of test/test_ghosts.jl with the default host generator.

C16, C8 = np.dtype(np.complex128), np.dtype(np.complex64)


def _panelbudget(panels, N, w, T):
    # test/setup.jl:146: whole 64-byte aligned panel blocks
    return panels * (-(-N * w * np.dtype(T).itemsize // 64)) * 64


def _blastol(N, k):
    return 100 * np.finfo(np.float64).eps * np.sqrt(N * k)


@pytest.mark.parametrize("w", (8, 7, 1))
def test_a_ghost_matrix_regenerates_the_same_entries(w):
    # test/test_ghosts.jl:1-17
    N, k = 24, 8
    O = F.GhostPanels(C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=w)
    assert O.shape == (N, k) and O.dtype == C16 and F.npanels(O) == -(-k // w)
    A = F.to_matrix(O)
    assert np.array_equal(A, F.to_matrix(O))
    assert np.array_equal(A, F.to_matrix(F.GhostPanels(C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=w)))
    assert not np.array_equal(A, F.to_matrix(F.GhostPanels(C16, N, k, plan=_plan(), seed=0xDEADBEEF + 1, w=w)))
    # Standard normals: nothing degenerate, independent columns.
    assert 0.5 < np.linalg.norm(A) / np.sqrt(N * k) < 2
    assert np.linalg.matrix_rank(A) == min(N, k)


def test_the_same_seed_gives_the_same_matrix_at_every_panel_width():
    # test/test_ghosts.jl:19-28
    N, k = 40, 23
    reference = F.to_matrix(F.GhostPanels(C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=23))
    for w in (23, 12, 7, 1):
        O = F.GhostPanels(C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=w)
        assert F.npanels(O) == -(-k // w)
        assert np.array_equal(F.to_matrix(O), reference)


def _normalized(dst, rng, col):
    F.randn_column(dst, rng, col)
    dst /= np.linalg.norm(dst)


@pytest.mark.parametrize("w", (23, 7, 1))
def test_a_per_column_generator_is_free_of_the_panel_boundaries(w):
    # test/test_ghosts.jl:40-56: normalizing a column needs the whole column,
    # and that is what the generator is handed however the panels cut it.
    N, k = 40, 23
    A = F.to_matrix(F.GhostPanels(_normalized, C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=w))
    assert np.allclose(np.linalg.norm(A, axis=0), 1.0, rtol=1e-14)
    assert np.array_equal(A, F.to_matrix(F.GhostPanels(_normalized, C16, N, k, plan=_plan(), seed=0xDEADBEEF, w=23)))


def test_a_rademacher_generator_round_trips():
    # test/test_ghosts.jl:58-68
    def rademacher(dst, rng, col):
        dst[...] = rng.choice(np.array([-1.0, 1.0]), size=dst.shape)

    O = F.GhostPanels(rademacher, C16, 24, 9, plan=_plan(), seed=7, w=4)
    A = F.to_matrix(O)
    assert np.all(np.abs(A) == 1)
    assert np.array_equal(A, F.to_matrix(O))


def test_a_narrowed_host_tier_hands_the_generator_its_own_eltype():
    # test/test_ghosts.jl:70-86: the generator writes into the host tier's
    # storage, a whole column however wide the panel holding it is.
    N, k, w = 12, 5, 2
    seen = []

    def f(dst, rng, col):
        seen.append((dst.dtype, dst.size))
        F.randn_column(dst, rng, col)

    F.to_matrix(F.GhostPanels(f, C16, N, k, plan=_plan(host_eltype=C8), w=w))
    assert len(seen) == k
    assert all(entry == (C8, N) for entry in seen)


def test_a_ghost_panel_is_regenerated_after_it_leaves_host_memory():
    # test/test_ghosts.jl:88-100: two blocks for four panels and no scratch
    # directory; there is nothing to write.
    N, k, w = 24, 8, 2
    plan = _plan(host_budget=_panelbudget(2, N, w, C16))
    O = F.GhostPanels(C16, N, k, plan=plan, seed=3, w=w)
    A = F.to_matrix(O)
    assert sum(1 for p in O.panels if P.isresident(p)) <= 2
    assert plan.scratch_dir is None
    assert np.array_equal(F.to_matrix(O), A)


@pytest.mark.parametrize("w", (4, 3))
def test_a_ghost_matrix_feeds_the_operator_sweeps(w):
    # test/test_ghosts.jl:102-122
    rng = np.random.default_rng(1)
    N, k = 20, 8
    backend = JitteryBackend(jitter=1e-4, seed=w)
    plan = _plan(backend)
    O = F.GhostPanels(C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = F.to_matrix(O)
    dense = rng.random((N, N)) - 0.5 + 1j * (rng.random((N, N)) - 0.5)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, F.DenseOperator(dense, backend), O)
    tol = _blastol(N, k)
    assert np.linalg.norm(F.to_matrix(Y) - dense @ A) <= tol * np.linalg.norm(dense @ A)
    S = A.conj().T @ (dense @ A)
    assert np.linalg.norm(F.gram(O, Y) - S) <= tol * np.linalg.norm(S)
    assert abs(F.norm(O) - np.linalg.norm(A)) <= tol * np.linalg.norm(A)
    seen = np.zeros((N, k), dtype=C16)

    def f(j, panel):
        r = F.panelrange(O, j)
        seen[:, r.start:r.stop] = backend.to_host(panel)

    F.foreachpanel(f, O, write=False)
    assert np.array_equal(seen, A)


def test_a_ghost_matrix_refuses_to_be_written_to():
    # test/test_ghosts.jl:124-140: refused before anything is staged
    N, k, w = 16, 6, 4
    backend = JitteryBackend(jitter=0.0)
    plan = _plan(backend)
    O = F.GhostPanels(C16, N, k, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    G = F.DenseOperator(np.eye(N, dtype=C16), backend)
    for attempt in (lambda: F.cholqr2(O), lambda: F.scale(O, 2.0), lambda: F.copyto(O, np.zeros((N, k), C16)),
                    lambda: F.panelmul(O, G, Y), lambda: F.axpy(1.0, Y, O),
                    lambda: F.foreachpanel(lambda j, panel: None, O)):
        with pytest.raises(F.ArgumentError):
            attempt()
    assert all(p.pins == 0 for p in O.panels)


def test_a_ghost_matrix_narrows_its_host_storage():
    # test/test_ghosts.jl:142-151
    N, k, w = 20, 7, 3
    O = F.GhostPanels(C16, N, k, plan=_plan(host_eltype=C8), seed=11, w=w)
    F.prefetch(O, 0)
    assert P.storage_eltype(O.panels[0]) == C8
    wide = F.to_matrix(F.GhostPanels(C16, N, k, plan=_plan(), seed=11, w=w))
    assert np.linalg.norm(F.to_matrix(O) - wide) <= 8 * np.finfo(np.float32).eps * np.linalg.norm(wide)


def test_ghost_panels_evict_without_writing_anything():
    # test/test_ghosts.jl:153-166
    plan = _plan()
    O = F.GhostPanels(C16, 16, 6, plan=plan, w=3)
    A = F.to_matrix(O)
    for panel in O.panels:
        assert P.isresident(panel)
        F.evict(panel, plan)
        assert not P.isresident(panel) and not P.ondisk(panel)
    assert np.array_equal(F.to_matrix(O), A)


@pytest.mark.parametrize("w", (23, 7, 1))
def test_a_read_only_sweep_leaves_a_ghost_regenerable(w):
    # test/test_ghosts.jl:168-197
    rng = np.random.default_rng(2)
    m, n, k = 40, 25, 23
    reference = rng.random((m, n)) - 0.5 + 1j * (rng.random((m, n)) - 0.5)
    backend = JitteryBackend(jitter=1e-4, seed=w)
    plan = _plan(backend)
    O = F.GhostPanels(C16, n, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = F.to_matrix(O)
    Y = F.PanelMatrix(m, k, plan=plan, w=w)
    F.panelmul(Y, F.DenseOperator(reference, backend), O)
    assert np.linalg.norm(F.to_matrix(Y) - reference @ A) <= _blastol(m, k) * np.linalg.norm(reference @ A)
    assert "ghost" in repr(O)
    assert not any(p.dirty for p in O.panels)
    assert all(p.pins == 0 for p in O.panels)
    for panel in O.panels:
        F.evict(panel, plan)
    assert np.array_equal(F.to_matrix(O), A)


def _sketchchain(plan, backend, reference, m, n, k, w):
    O = F.GhostPanels(C16, n, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y = F.PanelMatrix(m, k, plan=plan, w=w)
    F.panelmul(Y, F.DenseOperator(reference, backend), O)
    R = F.cholqr2(Y)
    return F.to_matrix(O), F.to_matrix(Y), R


def test_the_sketch_chain_matches_the_dense_reference_under_a_tight_budget(tmp_path):
    # test/test_ghosts.jl:199-218: the ghost gets fewer blocks than panels, so
    # its panels are dropped and made again while cholqr2's row sweeps hold
    # every panel of Y at once; then the same with a scratch directory.
    rng = np.random.default_rng(3)
    m, n, k, w = 40, 24, 24, 3
    reference = rng.random((m, n)) - 0.5 + 1j * (rng.random((m, n)) - 0.5)
    budget = _panelbudget(-(-k // w), m, w, C16) + _panelbudget(6, n, w, C16)
    backend = JitteryBackend(jitter=1e-4, seed=4)
    A, Q, R = _sketchchain(_plan(backend, host_budget=budget), backend, reference, m, n, k, w)
    bound = 100 * m * k * np.finfo(np.float64).eps
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(k), 2) <= bound
    assert np.linalg.norm(Q @ R - reference @ A) <= _blastol(m, k) * np.linalg.norm(reference @ A)
    spilled, Qd, Rd = _sketchchain(_plan(backend, host_budget=budget, scratch_dir=str(tmp_path)), backend,
                                   reference, m, n, k, w)
    assert np.array_equal(spilled, A)
    assert np.linalg.norm(Qd.conj().T @ Qd - np.eye(k), 2) <= bound
    assert np.linalg.norm(Qd @ Rd - reference @ spilled) <= _blastol(m, k) * np.linalg.norm(reference @ spilled)


def test_ghost_construction_is_validated():
    # test/test_ghosts.jl:220-227
    plan = _plan()
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(C16, 0, 4, plan=plan)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(C16, 4, 0, plan=plan)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(np.object_, 4, 4, plan=plan)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(C16, 4, 4, plan=plan, w=5)
    assert "ghost" in repr(F.GhostPanels(C16, 8, 4, plan=plan, w=2))
