"""cholqr2 by column panels (gram_panelwise, rdiv_panelwise,
cholqr2_panelwise) on the HIP backend against the oracle, including a
disk-backed basis larger than the host budget, which the row sweeps of
gram and cholqr2 refuse (test/test_disk.jl:224-234)."""
import numpy as np
import pytest
import scipy.linalg

import funicular_hip as F
from funicular_hip import plan as P
from oracle import ghost as OG
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

WIDTHS = (23, 12, 7, 1)
C16 = np.complex128


def randmatrix(rng, N, k, T=C16):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(T, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(T))


def make_plan(backend, **kw):
    kw.setdefault("device_budget", 64 * 2 ** 20)
    kw.setdefault("host_budget", 64 * 2 ** 20)
    return F.ResidencyPlan(backend=backend, **kw)


@pytest.mark.parametrize("T", [C16, np.complex64, np.float64])
@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_panelwise_against_the_oracle(backend, T, w, nb):
    print(f"panelwise T={np.dtype(T).name} w={w} nb={nb}", flush=True)
    rng = np.random.default_rng(10 * w + nb)
    N, k = 40, 23
    A = randmatrix(rng, N, k, T)
    plan = make_plan(backend)
    tol = OB.blastol(T, N, k)
    R, _ = OB.cholesky_upper(OB.gram(A), N, k)
    # group=1 is the pairwise schedule (one panel held); 3 leaves a ragged
    # last run; None holds as many panels as the device budget allows.
    for group in (1, 3, None):
        X = F.PanelMatrix.from_array(A, plan=plan, w=w)
        assert OB.isapprox(F.gram_panelwise(X, nbuffers=nb, group=group), OB.gram_panelwise(A, w), tol)
        F.rdiv_panelwise(X, R, nbuffers=nb, group=group)
        assert OB.isapprox(F.to_matrix(X), OB.rdiv_panelwise(A, R, w), tol)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Rq = F.cholqr2_panelwise(Y, nbuffers=nb)
    Qo, Ro = OB.cholqr2_panelwise(A, w)
    assert OB.isapprox(Rq, Ro, tol)
    Q = F.to_matrix(Y)
    assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(T, N, k)
    assert np.linalg.norm(Q @ Rq - A) <= tol * np.linalg.norm(A)


def test_panelwise_on_a_device_resident_plan(backend):
    rng = np.random.default_rng(3)
    N, k, w = 3000, 40, 8
    A = randmatrix(rng, N, k)
    plan = F.ResidencyPlan(backend=backend, device_budget=1 << 30, host_budget=1 << 30, device_resident=True)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R = F.cholqr2_panelwise(Y)
    Qo, Ro = OB.cholqr2(A, w)
    tol = OB.blastol(C16, N, k)
    assert OB.isapprox(R, Ro, tol)
    assert OB.isapprox(F.to_matrix(Y), Qo, tol)


def test_panelwise_survives_the_shifted_third_pass(backend):
    rng = np.random.default_rng(5)
    N, k, w = 400, 24, 5
    U = scipy.linalg.qr(randmatrix(rng, N, k), mode="economic")[0]
    V = scipy.linalg.qr(randmatrix(rng, k, k))[0]
    A = np.asfortranarray(U @ np.diag(10.0 ** np.linspace(0, -10, k)) @ V.conj().T)
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R = F.cholqr2_panelwise(Y)
    Q = F.to_matrix(Y)
    assert OB.orthogonality_error(Q) <= 1e-8
    assert np.linalg.norm(Q @ R - A) <= 1e-10 * np.linalg.norm(A)


def test_panelwise_cholqr2_of_a_basis_past_the_host_budget(backend, tmp_path):
    # Eight panels on disk behind a host budget of four: the row sweeps need
    # all eight host resident at once and refuse; the panel-wise passes only
    # ever stage a few.
    print("panelwise disk-backed", flush=True)
    rng = np.random.default_rng(9)
    N, k, w = 5000, 32, 4
    A = randmatrix(rng, N, k)
    budget = 4 * P.align_up(P.panel_bytes(N, w, np.dtype(C16)), P.SLAB_ALIGN)
    plan = F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=budget,
                           scratch_dir=str(tmp_path), panel_width=w)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    with pytest.raises(F.ArgumentError):
        F.gram(Y)
    assert all(p.pins == 0 for p in Y.panels)
    assert np.array_equal(F.to_matrix(Y), A)
    before = F.disk_reads(Y)
    R = F.cholqr2_panelwise(Y, group=3)
    assert F.disk_reads(Y) > before and F.disk_writes(Y) > 0
    assert all(p.pins == 0 for p in Y.panels)
    Qo, Ro = OB.cholqr2(A, w)
    tol = OB.blastol(C16, N, k)
    assert OB.isapprox(R, Ro, tol)
    assert OB.isapprox(F.to_matrix(Y), Qo, tol)


def test_panelwise_matches_the_row_sweeps_on_a_ghost_basis(backend):
    # N = 2^20: the two traversals sum in different orders, so they agree to
    # rounding, and both meet the orthogonality target.
    N, k, w = 1 << 20, 128, 16
    print(f"panelwise ghost N={N} k={k}", flush=True)
    plan = F.ResidencyPlan(backend=backend, device_budget=32 << 30, host_budget=32 << 30)
    O = F.GhostPanels(F.device_randn, C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y1 = F.PanelMatrix(N, k, plan=plan, w=w)
    Y2 = F.PanelMatrix(N, k, plan=plan, w=w)
    F.copyto(Y1, O)
    F.copyto(Y2, O)
    R1 = F.cholqr2(Y1)
    R2 = F.cholqr2_panelwise(Y2)
    assert np.linalg.norm(R1 - R2) <= 1e-12 * np.linalg.norm(R1)
    for Y in (Y1, Y2):
        G = F.gram(Y)
        assert np.linalg.norm(G - np.eye(k)) <= 1e-13
    for Y in (Y1, Y2):
        F.free(Y)


def test_panelwise_mixed_first_pass(backend, tmp_path):
    # The FP32 Gram and Cholesky only precondition: after the FP64 pass Q is
    # orthonormal to FP64 level and Q R reproduces Y to FP64 rounding (the
    # solve stays FP64), on a disk-backed basis past the host budget.
    rng = np.random.default_rng(13)
    N, k, w = 6000, 48, 8
    U = scipy.linalg.qr(randmatrix(rng, N, k), mode="economic")[0]
    V = scipy.linalg.qr(randmatrix(rng, k, k))[0]
    A = np.asfortranarray(U @ np.diag(10.0 ** np.linspace(0, -2, k)) @ V.conj().T)
    budget = 4 * P.align_up(P.panel_bytes(N, w, np.dtype(C16)), P.SLAB_ALIGN)
    plan = F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=budget,
                           scratch_dir=str(tmp_path), panel_width=w)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R = F.cholqr2_panelwise(Y, mixed=True)
    Q = F.to_matrix(Y)
    assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(C16, N, k)
    assert np.linalg.norm(Q @ R - A) <= OB.blastol(C16, N, k) * np.linalg.norm(A)
    Qo, Ro = OB.cholqr2(A, w)
    assert OB.isapprox(R, Ro, 1e-12)
    with pytest.raises(F.ArgumentError):
        F.cholqr2_panelwise(F.PanelMatrix(100, 4, plan=plan, w=2, dtype=np.complex64), mixed=True)
