"""Sweeps that stream panels through the disk tier on the HIP backend
(test/test_disk.jl is the model): panel sweeps, panelmul over a disk-backed
pair, and the row sweeps of gram and cholqr2, whose matrices cycle through
the scratch file while the producer thread reads ahead."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import plan as P
from funicular_hip.panels import ondisk
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

C16 = np.complex128


def randmatrix(rng, N, k, T=C16):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(T, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(T))


def panelbudget(n, N, w, T=C16):
    return n * P.align_up(P.panel_bytes(N, w, np.dtype(T)), P.SLAB_ALIGN)


def make_plan(backend, tmp_path, n, N, w, **kw):
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=panelbudget(n, N, w),
                           scratch_dir=str(tmp_path), panel_width=w, **kw)


@pytest.mark.parametrize("nb", (1, 2))
def test_panel_sweep_streams_through_disk(backend, tmp_path, nb):
    rng = np.random.default_rng(70)
    N, k, w = 1000, 24, 4
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, tmp_path, 5, N, w)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    before = F.disk_reads(pm)
    F.foreachpanel(lambda _, block: backend.axpby(block, block, 1.0, 1.0), pm, nbuffers=nb)
    assert F.disk_reads(pm) > before
    assert np.array_equal(F.to_matrix(pm), 2 * A)
    assert all(p.pins == 0 for p in pm.panels)


def test_panelmul_streams_a_disk_backed_pair(backend, tmp_path):
    rng = np.random.default_rng(71)
    N, k, w = 200, 24, 4
    A = randmatrix(rng, N, k)
    Gh = randmatrix(rng, N, N)
    # The host budget holds 8 panels; X and Y have 6 each.
    plan = make_plan(backend, tmp_path, 8, N, w)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, F.DenseOperator(Gh, backend), X)
    assert OB.isapprox(F.to_matrix(Y), Gh @ A, OB.blastol(C16, N, k))
    assert F.disk_reads(X) > 0
    assert F.disk_reads(Y) > 0 and F.disk_writes(Y) > 0


def test_row_sweeps_over_disk_backed_matrices(backend, tmp_path):
    rng = np.random.default_rng(72)
    N, k, w = 400, 24, 4
    A, B = randmatrix(rng, N, k), randmatrix(rng, N, k)
    # Row sweeps (gram, cholqr2) pin all 6 panels of a matrix at once; a
    # budget of 7 fits that for X or for Y, never for both together.
    plan = make_plan(backend, tmp_path, 7, N, w)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    R = F.cholqr2(Y)
    assert OB.isapprox(F.gram(X), A.conj().T @ A, OB.blastol(C16, N, k))
    assert F.disk_reads(X) > 0
    assert F.disk_writes(Y) > 0
    Q = F.to_matrix(Y)
    assert F.disk_reads(Y) > 0
    assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(C16, N, k)
    assert OB.isapprox(Q @ R, B, OB.blastol(C16, N, k))
    assert np.array_equal(F.to_matrix(X), A)


def test_cholqr2_on_a_loaded_file(backend, tmp_path):
    rng = np.random.default_rng(73)
    N, k, w = 3000, 16, 8
    A = randmatrix(rng, N, k)
    plan = F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20)
    path = str(tmp_path / "basis.panels")
    F.save(F.PanelMatrix.from_array(A, plan=plan, w=w), path)
    Y = F.load(path, plan=plan)
    assert all(ondisk(p) for p in Y.panels)
    R = F.cholqr2(Y)
    Qo, Ro = OB.cholqr2(A, w)
    assert OB.isapprox(R, Ro, 1e-12)
    assert OB.isapprox(F.to_matrix(Y), Qo, 1e-12)
    assert F.disk_reads(Y) == F.npanels(Y)


def test_narrowed_disk_tier_feeds_the_device(backend, tmp_path):
    rng = np.random.default_rng(74)
    N, k, w = 500, 16, 4
    A, B = randmatrix(rng, N, k), randmatrix(rng, N, k)
    S = np.complex64
    # Two matrices of four ComplexF32 panels in room for six: one cycles out.
    plan = F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=panelbudget(6, N, w, S),
                           scratch_dir=str(tmp_path), panel_width=w, host_eltype=S)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Z = F.PanelMatrix.from_array(B, plan=plan, w=w)
    assert X.store.stored == np.dtype(S) and X.store.computed == np.dtype(C16)
    A32 = A.astype(S).astype(C16)
    # Each staged row block is widened to ComplexF64 on the device.
    assert OB.isapprox(F.gram(X), A32.conj().T @ A32, 1e-12)
    assert F.disk_reads(X) > 0
    B32 = B.astype(S).astype(C16)
    assert OB.isapprox(F.gram(Z), B32.conj().T @ B32, 1e-12)


def test_failed_write_only_sweep_keeps_the_old_panels(backend, tmp_path):
    rng = np.random.default_rng(76)
    N, k, w = 300, 24, 4
    A, B = randmatrix(rng, N, k), randmatrix(rng, N, k)
    plan = make_plan(backend, tmp_path, 8, N, w)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    for p in Y.panels:
        F.evict(p, plan)

    def copy_then_fail(j, xb, yb):
        if j == 2:
            raise RuntimeError("panel function failed")
        backend.d2d(yb, xb)

    with pytest.raises(RuntimeError, match="panel function failed"):
        F.sweep(copy_then_fail, F.readpanels(X), F.writepanels(Y), nbuffers=2)
    assert all(p.pins == 0 for p in Y.panels)
    # Panels 0 and 1 were written; the rest hold what they held, including the
    # ones the producer had staged ahead of the failure.
    expect = B.copy()
    expect[:, :2 * w] = A[:, :2 * w]
    assert np.array_equal(F.to_matrix(Y), expect)
    assert np.array_equal(F.to_matrix(X), A)
