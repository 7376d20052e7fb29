"""Panel matrices, after test/test_panels.jl, on the CPU.

The host-side bookkeeping of PanelMatrix (shape, panel partition, narrowed
host storage, the host pool, eviction, copycols/copyto/similar, epochs) runs
on the jittered numpy backend (tests/fakes.py). Columns and panels are
numbered from 0 and column ranges are Python ranges.
"""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import panels as P
from funicular_hip import plan as PL

from fakes import JitteryBackend

PANEL_WIDTHS = (41, 21, 17, 1)  # test/test_panels.jl:1
C16, C8 = np.dtype(np.complex128), np.dtype(np.complex64)
EltTypes = (C16, np.dtype(np.float64), C8)


def randmatrix(rng, T, N, k):
    T = np.dtype(T)
    A = rng.standard_normal((N, k))
    if T.kind == "c":
        A = A + 1j * rng.standard_normal((N, k))
    return np.asfortranarray(A.astype(T))


def make_plan(backend=None, **kw):
    backend = backend if backend is not None else JitteryBackend(jitter=2e-4)
    kw.setdefault("host_budget", 64 << 20)
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, **kw)


def zerobuffer(plan, T, N, w):
    return plan.backend.todevice(np.zeros((N, w), dtype=T))


@pytest.mark.parametrize("w", PANEL_WIDTHS)
@pytest.mark.parametrize("T", EltTypes, ids=str)
def test_dense_round_trip(T, w):
    # test/test_panels.jl:3-23
    rng = np.random.default_rng(1)
    N, k = 13, 41
    A = randmatrix(rng, T, N, k)
    pm = F.PanelMatrix.from_array(A, plan=make_plan(), w=w)
    assert pm.shape == (N, k)
    assert pm.dtype == T
    assert F.panelwidth(pm) == w
    assert F.npanels(pm) == -(-k // w)
    assert sum(F.panelwidth(pm, j) for j in range(F.npanels(pm))) == k
    assert [c for j in range(F.npanels(pm)) for c in F.panelrange(pm, j)] == list(range(k))
    assert F.panelwidth(pm, F.npanels(pm) - 1) == (w if k % w == 0 else k % w)
    assert np.array_equal(F.to_matrix(pm), A)
    F.free(pm)


def test_ragged_last_panel_is_narrower():
    # test/test_panels.jl:25-35
    pm = F.PanelMatrix(8, 41, plan=make_plan(), w=17)
    assert F.npanels(pm) == 3
    assert [F.panelwidth(pm, j) for j in range(3)] == [17, 17, 7]
    assert P.panelstorage(pm.panels[2]).shape == (8, 7)
    assert F.panelrange(pm, 2) == range(34, 41)
    with pytest.raises(IndexError):
        F.panelrange(pm, 3)
    with pytest.raises(IndexError):
        F.panelwidth(pm, -1)


def test_narrowed_host_storage():
    # test/test_panels.jl:37-54
    rng = np.random.default_rng(2)
    N, k = 9, 20
    A = randmatrix(rng, C16, N, k)
    pm = F.PanelMatrix.from_array(A, plan=make_plan(host_eltype=C8), w=7)
    assert P.storage_eltype(pm.panels[0]) == C8
    assert pm.dtype == C16
    got = F.to_matrix(pm)
    assert np.array_equal(got, A.astype(C8).astype(C16))
    assert np.linalg.norm(got - A) <= 10 * np.finfo(np.float32).eps * np.linalg.norm(A)
    # A real host eltype would drop the imaginary parts of a complex matrix.
    with pytest.raises(F.ArgumentError):
        F.PanelMatrix.from_array(A, plan=make_plan(host_eltype=np.float32))
    real = F.PanelMatrix.from_array(randmatrix(rng, np.float64, N, k), plan=make_plan(host_eltype=np.float32))
    assert P.storage_eltype(real.panels[0]) == np.dtype(np.float32)


def test_no_scalar_indexing():
    # test/test_panels.jl:56-65: a PanelMatrix is not an array
    pm = F.PanelMatrix(4, 4, plan=make_plan())
    assert not isinstance(pm, np.ndarray)
    with pytest.raises(TypeError):
        pm[0, 0]
    with pytest.raises(TypeError):
        pm[0, 0] = 0
    with pytest.raises(TypeError):
        pm[2]


def test_constructor_validation():
    # test/test_panels.jl:67-74
    plan = make_plan()
    with pytest.raises(F.ArgumentError):
        F.PanelMatrix(0, 4, plan=plan)
    with pytest.raises(F.ArgumentError):
        F.PanelMatrix(4, 0, plan=plan)
    with pytest.raises(F.ArgumentError):
        F.PanelMatrix(4, 4, plan=plan, dtype=np.object_)
    with pytest.raises(F.DimensionMismatch):
        F.copyto(F.PanelMatrix(4, 4, plan=plan), np.zeros((4, 5), dtype=C16))


def test_conformality():
    # test/test_panels.jl:76-86
    plan = make_plan()
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    Y = F.PanelMatrix(10, 20, plan=plan, w=7)
    Z = F.PanelMatrix(10, 20, plan=plan, w=5)
    W = F.PanelMatrix(11, 20, plan=plan, w=7)
    assert P.assert_conformal(X, Y) is None
    with pytest.raises(F.ArgumentError):
        P.assert_conformal(X, Z)
    with pytest.raises(F.ArgumentError):
        P.assert_conformal(X, W)


def test_collect_guard():
    # test/test_panels.jl:88-94
    pm = F.PanelMatrix(16, 16, plan=make_plan())
    with pytest.raises(F.ArgumentError):
        F.to_matrix(pm, max_bytes=16 * 16 * 16 - 1)
    assert F.to_matrix(pm, max_bytes=16 * 16 * 16).shape == (16, 16)


def test_show():
    # test/test_panels.jl:96-111
    pm = F.PanelMatrix(100, 41, plan=make_plan(), w=17)
    text = repr(pm)
    assert "100×41" in text and "w = 17" in text and "3 panels" in text
    assert "host storage" not in text
    narrow = F.PanelMatrix(100, 41, plan=make_plan(host_eltype=C8), w=17)
    assert "host storage complex64" in repr(narrow)
    F.free(narrow)
    assert "released" in repr(narrow)


def test_release_and_eviction():
    # test/test_panels.jl:113-133
    plan = make_plan(host_budget=1 << 20)
    pm = F.PanelMatrix(1024, 8, plan=plan, w=4)
    used = PL.host_bytes_in_use(plan.hostpool)
    assert used == 2 * 1024 * 4 * 16
    # No colder tier: neither a clean nor a dirty panel can be evicted.
    with pytest.raises(F.ArgumentError):
        F.evict(pm.panels[0], plan)
    pm.panels[0].dirty = True
    with pytest.raises(F.ArgumentError):
        F.evict(pm.panels[0], plan)
    F.free(pm)
    assert PL.host_bytes_in_use(plan.hostpool) == 0
    assert not P.isresident(pm.panels[0])
    with pytest.raises(F.ArgumentError):
        P.materialize(zerobuffer(plan, C16, 1024, 4), pm.panels[0], plan)
    with pytest.raises(F.ArgumentError):
        P.writeback(pm.panels[0], zerobuffer(plan, C16, 1024, 4), plan)
    F.PanelMatrix(1024, 8, plan=plan, w=4)
    assert PL.host_bytes_in_use(plan.hostpool) == used
    assert PL.host_bytes_reserved(plan.hostpool) <= plan.hostpool.budget


def test_failed_construction_releases_its_blocks():
    # test/test_panels.jl:135-142
    plan = make_plan(host_budget=3 * 1024 * 4 * 16)
    pool = plan.hostpool
    with pytest.raises(F.ArgumentError):
        F.PanelMatrix(1024, 16, plan=plan, w=4)
    assert PL.host_bytes_in_use(pool) == 0
    assert PL.host_bytes_free(pool) == sum(pool.cursors)


@pytest.mark.parametrize("w", PANEL_WIDTHS)
@pytest.mark.parametrize("T", EltTypes, ids=str)
def test_copycols_from_a_host_matrix(T, w):
    # test/test_panels.jl:144-159
    rng = np.random.default_rng(3)
    N, k = 13, 41
    A = randmatrix(rng, T, N, k)
    plan = make_plan()
    for dcols in (range(0, 9), range(11, 28), range(32, 41), range(0, 41), range(19, 20)):
        dest = F.PanelMatrix.from_array(A, plan=plan, w=w)
        B = randmatrix(rng, T, N, len(dcols))
        reference = A.copy()
        reference[:, dcols.start:dcols.stop] = B
        assert F.copycols(dest, dcols, B) is dest
        assert np.array_equal(F.to_matrix(dest), reference)
        assert all(p.pins == 0 for p in dest.panels)
        F.free(dest)


@pytest.mark.parametrize("widths", ((7, 7), (7, 5), (23, 4), (1, 6)))
@pytest.mark.parametrize("T", EltTypes, ids=str)
def test_copycols_between_panel_matrices(T, widths):
    # test/test_panels.jl:161-181
    sw, dw = widths
    rng = np.random.default_rng(4)
    N, sk, dk = 12, 23, 30
    A = randmatrix(rng, T, N, sk)
    B = randmatrix(rng, T, N, dk)
    plan = make_plan()
    for dcols, scols in ((range(0, sk), range(0, sk)), (range(7, 20), range(2, 15)), (range(24, 30), range(0, 6)),
                         (range(13, 14), range(22, 23))):
        src = F.PanelMatrix.from_array(A, plan=plan, w=sw)
        dest = F.PanelMatrix.from_array(B, plan=plan, w=dw)
        reference = B.copy()
        reference[:, dcols.start:dcols.stop] = A[:, scols.start:scols.stop]
        assert F.copycols(dest, dcols, src, scols) is dest
        assert np.array_equal(F.to_matrix(dest), reference)
        assert np.array_equal(F.to_matrix(src), A)
        assert all(p.pins == 0 for p in src.panels + dest.panels)
        F.free(src)
        F.free(dest)


def test_copycols_grows_a_sketch():
    # test/test_panels.jl:183-194
    rng = np.random.default_rng(5)
    N, m = 20, 9
    plan = make_plan()
    A = randmatrix(rng, C16, N, m)
    narrow = F.PanelMatrix.from_array(A, plan=plan, w=4)
    wide = F.PanelMatrix(N, 2 * m, plan=plan, w=5)
    F.copycols(wide, range(0, m), narrow, range(0, m))
    B = randmatrix(rng, C16, N, m)
    F.copycols(wide, range(m, 2 * m), B)
    got = F.to_matrix(wide)
    assert np.array_equal(got[:, :m], A)
    assert np.array_equal(got[:, m:], B)


def test_a_ghost_matrix_is_a_copycols_source():
    # test/test_panels.jl:196-213
    N, k, w = 20, 8, 3
    plan = make_plan()
    omega = F.GhostPanels(C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = F.to_matrix(omega)
    dest = F.PanelMatrix(N, 2 * k, plan=plan, w=5)
    F.copycols(dest, range(0, k), omega, range(0, k))
    F.copycols(dest, range(k, 2 * k), omega, range(0, k))
    got = F.to_matrix(dest)
    assert np.array_equal(got[:, :k], A)
    assert np.array_equal(got[:, k:], A)
    # Reading a ghost leaves it a ghost, with nothing written and nothing dirty.
    assert "ghost" in repr(omega)
    assert not any(p.dirty for p in omega.panels)
    assert np.array_equal(F.to_matrix(omega), A)


def test_copyto_repartitions_a_matrix():
    # test/test_panels.jl:215-228
    rng = np.random.default_rng(6)
    N, k = 9, 20
    A = randmatrix(rng, C16, N, k)
    plan = make_plan()
    src = F.PanelMatrix.from_array(A, plan=plan, w=7)
    dest = F.PanelMatrix(N, k, plan=plan, w=4)
    assert F.copyto(dest, src) is dest
    assert F.panelwidth(dest) == 4
    assert np.array_equal(F.to_matrix(dest), A)
    with pytest.raises(F.DimensionMismatch):
        F.copyto(F.PanelMatrix(N, k + 1, plan=plan, w=4), src)
    with pytest.raises(F.DimensionMismatch):
        F.copyto(F.PanelMatrix(N + 1, k, plan=plan, w=4), src)


def test_copycols_through_narrowed_host_storage():
    # test/test_panels.jl:230-244
    rng = np.random.default_rng(7)
    N, k = 9, 12
    A = randmatrix(rng, C16, N, k)
    narrowed = A.astype(C8).astype(C16)
    src = F.PanelMatrix.from_array(A, plan=make_plan(host_eltype=C8), w=5)
    dest = F.PanelMatrix(N, k, plan=make_plan(), w=4)
    F.copycols(dest, range(0, k), src, range(0, k))
    assert np.array_equal(F.to_matrix(dest), narrowed)
    stored = F.PanelMatrix(N, k, plan=make_plan(host_eltype=C8), w=3)
    F.copycols(stored, range(0, k), A)
    assert np.array_equal(F.to_matrix(stored), narrowed)


def test_copycols_validation():
    # test/test_panels.jl:246-268
    rng = np.random.default_rng(8)
    N, k = 8, 12
    plan = make_plan()
    dest = F.PanelMatrix(N, k, plan=plan, w=5)
    src = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=4)
    A = randmatrix(rng, C16, N, 4)
    omega = F.GhostPanels(C16, N, k, plan=plan, w=4)
    with pytest.raises(F.ArgumentError):
        F.copycols(dest, range(9, 14), A)
    with pytest.raises(F.ArgumentError):
        F.copycols(dest, range(-1, 3), A)
    with pytest.raises(F.DimensionMismatch):
        F.copycols(dest, range(0, 4), randmatrix(rng, C16, N, 5))
    with pytest.raises(F.DimensionMismatch):
        F.copycols(dest, range(0, 4), randmatrix(rng, C16, N + 1, 4))
    with pytest.raises(F.ArgumentError):
        F.copycols(dest, range(0, 4), src, range(9, 13))
    with pytest.raises(F.DimensionMismatch):
        F.copycols(dest, range(0, 4), src, range(0, 5))
    with pytest.raises(F.DimensionMismatch):
        F.copycols(dest, range(0, 4), F.PanelMatrix(N + 1, k, plan=plan, w=4), range(0, 4))
    with pytest.raises(F.ArgumentError):
        F.copycols(dest, range(0, 4), dest, range(4, 8))
    with pytest.raises(F.ArgumentError):
        F.copycols(omega, range(0, 4), A)
    with pytest.raises(F.ArgumentError):
        F.copycols(omega, range(0, 4), src, range(0, 4))
    # Every one of those is refused before anything is staged.
    assert all(p.pins == 0 for p in dest.panels + src.panels)


def test_similar_takes_the_plan_and_the_width_along():
    # test/test_panels.jl:270-293
    N, k = 40, 23
    plan = make_plan()
    pm = F.PanelMatrix(N, k, plan=plan, w=7)
    like = F.similar(pm)
    assert like.shape == (N, k) and like.dtype == C16 and F.panelwidth(like) == 7
    assert like.plan is plan
    S = np.dtype(np.float32)
    assert F.similar(pm, S).dtype == S
    assert F.similar(pm, S).shape == (N, k)
    assert F.panelwidth(F.similar(pm, S)) == 7
    tall = F.similar(pm, C16, (N + 17, k))
    assert tall.shape == (N + 17, k) and F.panelwidth(tall) == 7
    # A width wider than the matrix would leave the first panel ragged.
    assert F.panelwidth(F.similar(pm, C16, (N, 3))) == 3


def test_a_similar_companion_is_panelmul_conformal():
    # test/test_panels.jl:295-307
    rng = np.random.default_rng(9)
    N, m, k = 40, 17, 23
    backend = JitteryBackend(jitter=2e-4)
    plan = make_plan(backend)
    A = randmatrix(rng, C16, N, k)
    G = randmatrix(rng, C16, m, N)
    X = F.PanelMatrix.from_array(A, plan=plan, w=7)
    Y = F.similar(X, C16, (m, k))
    F.panelmul(Y, F.DenseOperator(G, backend), X)
    assert np.linalg.norm(F.to_matrix(Y) - G @ A) <= 1e-13 * np.linalg.norm(G @ A)


def test_copycols_stages_both_sides_through_the_disk_tier(tmp_path):
    # test/test_panels.jl:309-331: room for the two panels one run holds and
    # no more, so every other panel of both matrices is on disk meanwhile.
    rng = np.random.default_rng(10)
    N, k, w = 10, 24, 4
    A = randmatrix(rng, C16, N, k)
    B = randmatrix(rng, C16, N, k)
    plan = make_plan(scratch_dir=str(tmp_path), panel_width=w, host_budget=2 * PL.panel_bytes(N, w, C16))
    src = F.PanelMatrix.from_array(A, plan=plan, w=w)
    dest = F.PanelMatrix.from_array(B, plan=plan, w=w)
    reference = B.copy()
    reference[:, 4:20] = A[:, 0:16]
    F.copycols(dest, range(4, 20), src, range(0, 16))
    assert np.array_equal(F.to_matrix(dest), reference)
    assert np.array_equal(F.to_matrix(src), A)
    assert F.disk_reads(src) > 0 and F.disk_reads(dest) > 0 and F.disk_writes(dest) > 0


def test_similar_opens_a_store_of_its_own(tmp_path):
    # test/test_panels.jl:333-347
    plan = make_plan(scratch_dir=str(tmp_path), panel_width=4)
    pm = F.PanelMatrix(8, 12, plan=plan, w=4)
    assert len(list(tmp_path.iterdir())) == 1
    like = F.similar(pm)
    assert len(list(tmp_path.iterdir())) == 2
    assert like.store is not pm.store
    F.free(like)
    assert len(list(tmp_path.iterdir())) == 1
    F.free(pm)


def test_epoch_and_dirty_bookkeeping():
    # test/test_panels.jl:350-365
    rng = np.random.default_rng(11)
    plan = make_plan(JitteryBackend(jitter=0.0))
    pm = F.PanelMatrix(6, 6, plan=plan, w=3)
    assert all(p.epoch == 0 and not p.dirty for p in pm.panels)
    F.copyto(pm, randmatrix(rng, C16, 6, 6))
    assert all(p.epoch == 1 and p.dirty for p in pm.panels)
    P.writeback(pm.panels[0], zerobuffer(plan, C16, 6, 3), plan)
    assert pm.panels[0].epoch == 2
    assert pm.panels[1].epoch == 1
    P.materialize(zerobuffer(plan, C16, 6, 3), pm.panels[1], plan)
    assert pm.panels[1].epoch == 1
