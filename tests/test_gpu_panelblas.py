"""Panel BLAS and the sweep pipeline on the HIP backend against the oracle
(test/test_panelblas.jl, test/test_executor.jl, test/test_ghosts.jl are the
models)."""
import numpy as np
import pytest
import scipy.linalg

import funicular_hip as F
from funicular_hip import plan as P
from funicular_hip.executor import PanelSteps, RowSteps, readpanels, rowtraversal, sweep, updatepanels, writepanels
from oracle import ghost as OG
from oracle import operators as OO
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

WIDTHS = (23, 12, 7, 1)
C16 = np.complex128


def randmatrix(rng, N, k, dtype=C16):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(dtype, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(dtype))


def make_plan(backend, **kw):
    kw.setdefault("device_budget", 64 * 2 ** 20)
    kw.setdefault("host_budget", 64 * 2 ** 20)
    return F.ResidencyPlan(backend=backend, **kw)


@pytest.mark.parametrize("T", [C16, np.complex64, np.float64])
@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_gram_against_dense(backend, T, w, nb):
    rng = np.random.default_rng(w * 10 + nb)
    N, k = 40, 23
    A, B = randmatrix(rng, N, k, T), randmatrix(rng, N, k, T)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    tol = OB.blastol(T, N, k)
    assert OB.isapprox(F.gram(X, Y, nbuffers=nb), OB.gram(A, B, w), tol)
    assert OB.isapprox(F.gram(X, nbuffers=nb), OB.gram(A, w=w), tol)
    assert np.array_equal(F.to_matrix(X), A)


@pytest.mark.parametrize("w", WIDTHS)
def test_gram_same_bits_at_nbuffers_1_and_2(backend, w):
    rng = np.random.default_rng(w)
    A, B = randmatrix(rng, 40, 23), randmatrix(rng, 40, 23)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    assert np.array_equal(F.gram(X, Y, nbuffers=1), F.gram(X, Y, nbuffers=2))
    assert np.array_equal(F.gram(X, nbuffers=1), F.gram(X, nbuffers=2))


def test_gram_rejects_mismatched_operands(backend):
    plan = make_plan(backend)
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=plan, w=5))
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=plan, w=7, dtype=np.complex64))
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=make_plan(backend), w=7))


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_cholqr2_orthonormalizes(backend, w, nb):
    rng = np.random.default_rng(100 + w)
    N, k = 40, 23
    A = randmatrix(rng, N, k)
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R = F.cholqr2(Y, nbuffers=nb)
    Q = F.to_matrix(Y)
    assert R.shape == (k, k)
    assert np.allclose(R, np.triu(R))
    assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(C16, N, k)
    assert OB.isapprox(Q @ R, A, OB.blastol(C16, N, k))
    Qo, Ro = OB.cholqr2(A, w)
    assert OB.isapprox(Q, Qo, 1e-12) and OB.isapprox(R, Ro, 1e-12)


@pytest.mark.parametrize("kappa", (1e2, 1e6, 1e10))
def test_cholqr2_survives_conditioning(backend, kappa):
    # test/test_panelblas.jl:284-325
    rng = np.random.default_rng(int(np.log10(kappa)))
    N, k, w = 40, 23, 7
    U = scipy.linalg.qr(randmatrix(rng, N, k), mode="economic")[0]
    V = scipy.linalg.qr(randmatrix(rng, k, k))[0]
    A = np.asfortranarray(U @ np.diag(10.0 ** np.linspace(0, -np.log10(kappa), k)) @ V.conj().T)
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R = F.cholqr2(Y)
    Q = F.to_matrix(Y)
    shifted = kappa > 1 / np.sqrt(np.finfo(float).eps)
    bound = 1e-8 if shifted else OB.orthogonality_bound(C16, N, k)
    assert OB.orthogonality_error(Q) <= bound
    assert np.linalg.norm(Q @ R - A) <= 1e-10 * np.linalg.norm(A)


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_project_and_panelmul_against_dense(backend, w, nb):
    rng = np.random.default_rng(200 + w)
    N, k = 40, 23
    Gh = randmatrix(rng, N, N)
    G = F.DenseOperator(Gh, backend)
    A = randmatrix(rng, N, k)
    plan = make_plan(backend)
    Q = F.PanelMatrix.from_array(A, plan=plan, w=w)
    tol = OB.blastol(C16, N, k)
    S = F.project(Q, G, nbuffers=nb)
    assert OB.isapprox(S, OB.project(A, lambda X: Gh @ X, w), tol)
    assert np.array_equal(F.to_matrix(Q), A)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, Q, nbuffers=nb)
    assert OB.isapprox(F.to_matrix(Y), OB.panelmul(lambda X: Gh @ X, A, w), tol)
    Z = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Z, G.adjoint(), Q, nbuffers=nb)
    assert OB.isapprox(F.to_matrix(Z), Gh.conj().T @ A, tol)


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_rightmul_against_dense(backend, w, nb):
    rng = np.random.default_rng(300 + w)
    N, s = 40, 23
    A = randmatrix(rng, N, s)
    C = randmatrix(rng, s, 12)
    D = randmatrix(rng, s, s)
    plan = make_plan(backend)
    src = F.PanelMatrix.from_array(A, plan=plan, w=w)
    dest = F.PanelMatrix(N, 12, plan=plan, w=max(1, min(w - 3, 12)))
    assert F.rightmul(dest, src, C, nbuffers=nb) is dest
    assert OB.isapprox(F.to_matrix(dest), A @ C, OB.blastol(C16, N, s))
    assert F.rightmul(src, D, nbuffers=nb) is src
    assert OB.isapprox(F.to_matrix(src), A @ D, OB.blastol(C16, N, s))


def test_rightmul_same_bits_at_nbuffers(backend):
    rng = np.random.default_rng(301)
    A, C = randmatrix(rng, 40, 23), randmatrix(rng, 23, 12)
    out = []
    for nb in (1, 2):
        plan = make_plan(backend)
        src = F.PanelMatrix.from_array(A, plan=plan, w=7)
        dest = F.PanelMatrix(40, 12, plan=plan, w=4)
        out.append(F.to_matrix(F.rightmul(dest, src, C, nbuffers=nb)))
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("w", (8, 7, 1))
def test_ghost_regenerates_and_matches_oracle(backend, w):
    N, k = 24, 8
    plan = make_plan(backend)
    O = F.GhostPanels(F.device_randn, C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = F.to_matrix(O)
    assert np.array_equal(A, OG.matrix(C16, N, k, 0xDEADBEEF))
    assert np.array_equal(A, F.to_matrix(F.GhostPanels(F.device_randn, C16, N, k, plan=make_plan(backend), seed=0xDEADBEEF, w=3)))
    assert not np.array_equal(A, F.to_matrix(F.GhostPanels(F.device_randn, C16, N, k, plan=plan, seed=0xDEADBEEF + 1, w=w)))
    assert 0.5 < np.linalg.norm(A) / np.sqrt(N * k) < 2
    assert np.linalg.matrix_rank(A) == min(N, k)


def test_ghost_feeds_sweeps_and_refuses_writes(backend):
    rng = np.random.default_rng(400)
    N, k, w = 40, 23, 7
    plan = make_plan(backend)
    O = F.GhostPanels(F.device_randn, C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = OG.matrix(C16, N, k, 0xDEADBEEF)
    Gh = randmatrix(rng, N, N)
    G = F.DenseOperator(Gh, backend)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, O)
    assert OB.isapprox(F.to_matrix(Y), Gh @ A, OB.blastol(C16, N, k))
    assert OB.isapprox(F.gram(O, Y), A.conj().T @ Gh @ A, OB.blastol(C16, N, k))
    C = randmatrix(rng, k, 12)
    dest = F.PanelMatrix(N, 12, plan=plan, w=4)
    F.rightmul(dest, O, C)
    assert OB.isapprox(F.to_matrix(dest), A @ C, OB.blastol(C16, N, k))
    with pytest.raises(F.ArgumentError):
        F.cholqr2(O)
    with pytest.raises(F.ArgumentError):
        F.panelmul(O, G, Y)
    assert all(p.pins == 0 for p in O.panels)


def test_sketch_chain_matches_the_oracle(backend):
    # The 1-D sketch chain at a small N: ghost -> tridiagonal -> cholqr2 -> project.
    N, k, w = 4096, 64, 8
    plan = make_plan(backend)
    O = F.GhostPanels(F.device_randn, C16, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, backend=backend)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, O)
    R = F.cholqr2(Y)
    S = F.project(Y, G)
    Om = OG.matrix(C16, N, k, 0xDEADBEEF)
    Yo = OB.panelmul(lambda X: OO.tridiag_apply(X, 2 + 0.1j, -1.0), Om, w)
    Qo, Ro = OB.cholqr2(Yo, w)
    So = OB.project(Qo, lambda X: OO.tridiag_apply(X, 2 + 0.1j, -1.0), w)
    Q = F.to_matrix(Y)
    assert OB.orthogonality_error(Q) <= 1e-13
    assert np.linalg.norm(Q @ R - Yo) / np.linalg.norm(Yo) <= 1e-13
    assert OB.isapprox(S, So, 1e-13)


@pytest.mark.parametrize("nb", (1, 2))
def test_device_resident_plan(backend, nb):
    rng = np.random.default_rng(500)
    N, k, w = 3000, 96, 16
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, device_budget=256 * 2 ** 20, device_resident=True)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    assert Y.resident
    assert OB.isapprox(F.gram(Y, nbuffers=nb), OB.gram(A, w=w), OB.blastol(C16, N, k))
    R = F.cholqr2(Y, nbuffers=nb)
    Q = F.to_matrix(Y)
    assert OB.orthogonality_error(Q) <= 1e-13
    assert np.linalg.norm(Q @ R - A) / np.linalg.norm(A) <= 1e-13
    # Same bits as the streamed path: the kernels see the same row blocks.
    hp = make_plan(backend, device_budget=256 * 2 ** 20)
    Yh = F.PanelMatrix.from_array(A, plan=hp, w=w)
    Rh = F.cholqr2(Yh, nbuffers=nb)
    assert np.array_equal(R, Rh)
    assert np.array_equal(Q, F.to_matrix(Yh))


@pytest.mark.parametrize("mixed", (False, True))
def test_fused_schedule_uploads_y_three_times(backend, mixed):
    # On the fused schedule, off HBM, each solve sweep also forms the next
    # pass's Gram, so cholqr2 uploads Y three times (Gram, solve + Gram,
    # solve) instead of the reference's four, and the bits match the resident
    # run, whose passes are not fused.
    rng = np.random.default_rng(501)
    N, k, w = 3000, 96, 16
    A = randmatrix(rng, N, k)
    run = ((lambda Y: F.cholqr2_mixed(Y, schedule="fused")) if mixed
           else (lambda Y: F.cholqr2(Y, schedule="fused")))
    hp = make_plan(backend, device_budget=256 * 2 ** 20)
    Yh = F.PanelMatrix.from_array(A, plan=hp, w=w)
    c0 = F._lib.counters()
    Rh = run(Yh)
    c1 = F._lib.counters()
    up = c1["h2d_bytes"] - c0["h2d_bytes"]
    assert 3 * N * k * 16 <= up < 3 * N * k * 16 + 8 * k * k * 16, up
    rp = make_plan(backend, device_budget=256 * 2 ** 20, device_resident=True)
    Yr = F.PanelMatrix.from_array(A, plan=rp, w=w)
    Rr = run(Yr)
    assert np.array_equal(Rh, Rr)
    assert np.array_equal(F.to_matrix(Yh), F.to_matrix(Yr))
    assert OB.orthogonality_error(F.to_matrix(Yh)) <= 1e-13


def test_fused_passes_round_to_a_narrower_host_tier(backend):
    # With ComplexF32 host storage the fused Gram is formed after rounding
    # each solved block to ComplexF32, so cholqr2 matches the unfused passes
    # (a pass, then a pass on what was stored) to the bit.
    from funicular_hip.panelblas import cholqr_pass, cholqr_pass_then
    rng = np.random.default_rng(502)
    N, k, w = 3000, 64, 16
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, device_budget=256 * 2 ** 20, host_eltype=np.complex64)
    Y1 = F.PanelMatrix.from_array(A, plan=plan, w=w)
    c0 = F._lib.counters()
    R1 = F.cholqr2(Y1, schedule="fused")
    c1 = F._lib.counters()
    up = c1["h2d_bytes"] - c0["h2d_bytes"]
    assert 3 * N * k * 8 <= up < 3 * N * k * 8 + 8 * k * k * 16, up
    Y2 = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R2, s1 = cholqr_pass(Y2, schedule="fused")
    R2, s2 = cholqr_pass_then(Y2, R2)
    assert not (s1 or s2)
    assert np.array_equal(R1, R2)
    assert np.array_equal(F.to_matrix(Y1), F.to_matrix(Y2))


@pytest.mark.parametrize("w", (23, 8, 1))
def test_project_on_a_device_resident_plan(backend, w):
    # On the fused schedule, with room for a second N x k buffer, G Q is
    # formed whole and S is one Gram sweep; with no such room, the
    # reference's panel schedule runs.
    print(f"resident project w={w}", flush=True)
    rng = np.random.default_rng(700 + w)
    N, k = 20000, 64
    A = randmatrix(rng, N, k)
    tol = OB.blastol(C16, N, k)
    d, o = 2 + 0.1j, -1.0
    ref = OB.project(A, lambda X: OO.tridiag_apply(X, d, o), w)
    matrix = N * k * 16
    # The plan needs two N x w buffers besides the matrix; w < k / 2 leaves
    # the second budget short of a second N x k.
    for budget, pooled in ((4 * matrix, True), (matrix + max(2 * N * w * 16, 10 << 20), False)):
        plan = make_plan(backend, device_budget=budget, device_resident=True)
        G = F.Tridiagonal1D(N, d, o, backend=backend)
        Q = F.PanelMatrix.from_array(A, plan=plan, w=w)
        S = F.project(Q, G, schedule="fused")
        assert OB.isapprox(S, ref, tol)
        assert np.array_equal(F.to_matrix(Q), A)
        # The pooled route leaves its N x k buffer in the pool for reuse.
        assert bool(plan.devicepool.available.get((np.dtype(C16), (N, k)))) == pooled


@pytest.mark.parametrize("w", (23, 8, 7, 1))
@pytest.mark.parametrize("nb", (1, 2))
def test_sweeps_move_every_panel(backend, w, nb):
    rng = np.random.default_rng(600 + w)
    N, k = 9, 23
    A = randmatrix(rng, N, k)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    order = []
    seen = {}

    def look(j, panel):
        order.append(j)
        seen[j] = panel.shape
    F.foreachpanel(look, X, write=False, nbuffers=nb)
    assert order == list(range(F.npanels(X)))
    assert all(seen[j] == (N, F.panelwidth(X, j)) for j in seen)

    def scale(j, src, dst):
        backend.convert_copy(dst, src)
    sweep(scale, readpanels(X), writepanels(Y), nbuffers=nb)
    assert np.array_equal(F.to_matrix(Y), A)
    assert all(p.epoch == 1 and p.dirty for p in Y.panels)
    # Building X from an array wrote each panel once; a read sweep leaves that.
    assert all(p.epoch == 1 for p in X.panels)
    before = P.device_bytes_allocated(plan)
    sweep(scale, readpanels(X), writepanels(Y), nbuffers=nb)
    assert P.device_bytes_allocated(plan) == before


def test_row_sweep_and_failure_path(backend):
    rng = np.random.default_rng(700)
    N, k, w = 13, 23, 7
    A = randmatrix(rng, N, k)
    R = np.asfortranarray(np.triu(randmatrix(rng, k, k)) + 4 * np.eye(k))
    plan = make_plan(backend)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    factor = backend.todevice(R)
    sweep(lambda _, b: backend.rdiv_upper(b, factor), updatepanels(pm), traversal=rowtraversal(pm))
    assert OB.isapprox(F.to_matrix(pm), OB.rdiv_upper(A, R), OB.blastol(C16, N, k))
    F.foreachpanel(lambda j, panel: None, pm)
    before = P.device_bytes_allocated(plan)

    def boom(j, panel):
        if j == 1:
            raise RuntimeError("boom")
    with pytest.raises(RuntimeError):
        F.foreachpanel(boom, pm)
    assert P.device_bytes_allocated(plan) == before
    assert all(p.pins == 0 for p in pm.panels)
    with pytest.raises(F.ArgumentError):
        sweep(lambda *a: None, updatepanels(pm), traversal=PanelSteps([0, 1, 0], [0, 1, 2]))
    sweep(lambda *a: None, readpanels(pm), traversal=PanelSteps([0, 1, 0], [0, 1, 2]))


def test_narrowed_host_storage(backend):
    rng = np.random.default_rng(800)
    N, k, w = 40, 23, 7
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, host_eltype=np.complex64)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    stored = A.astype(np.complex64).astype(C16)
    assert OB.isapprox(F.gram(X), stored.conj().T @ stored, OB.blastol(C16, N, k))


@pytest.mark.parametrize("nb", (1, 2, 3))
@pytest.mark.parametrize("narrow", (False, True))
def test_row_sweep_with_read_and_updated_operands(backend, nb, narrow):
    """A row-block sweep that reads one host matrix and updates another: each
    upload into a slot waits for the compute that last read it (read-only
    operand) or, panel by panel, for the write-back that last read its columns
    (updated operand). Every step's result must land exactly, at every
    nbuffers, with the updated blocks narrowed to ComplexF32 on the host too."""
    rng = np.random.default_rng(900 + nb)
    N, k, w = 3000, 96, 4
    A, B = randmatrix(rng, N, k), randmatrix(rng, N, k)
    kw = {"host_eltype": np.complex64} if narrow else {}
    plan = make_plan(backend, **kw)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    t = rowtraversal(Y)
    assert len(t.blocks) > 3 * nb

    def f(_, xb, yb):
        backend.axpby(yb, xb, 0.5, 2.0)  # yb <- 0.5 xb + 2 yb

    sweep(f, readpanels(X), updatepanels(Y), nbuffers=nb, traversal=t)
    sweep(f, readpanels(X), updatepanels(Y), nbuffers=nb, traversal=t)
    if narrow:
        A, B = (M.astype(np.complex64).astype(C16) for M in (A, B))
        once = (0.5 * A + 2 * B).astype(np.complex64).astype(C16)
        expect = (0.5 * A + 2 * once).astype(np.complex64).astype(C16)
        assert np.allclose(F.to_matrix(Y), expect, rtol=0, atol=1e-6)
    else:
        expect = 0.5 * A + 2 * (0.5 * A + 2 * B)
        assert np.allclose(F.to_matrix(Y), expect, rtol=1e-15, atol=1e-15)
    assert np.array_equal(F.to_matrix(X), A)
    assert all(p.pins == 0 for p in X.panels + Y.panels)


@pytest.mark.parametrize("group", (1, 2, 3, None))
def test_project_by_runs_of_held_images(backend, group):
    """Off HBM, project holds the images of a run of panels on the device and
    streams every panel past each run once: (1 + number of runs) uploads of the
    basis, against p + 1 for the reference schedule (group 1, which it runs
    unchanged). S agrees with the oracle at every run length, ragged last
    panel and run included."""
    rng = np.random.default_rng(950 + (group or 0))
    N, k, w = 3000, 23, 4
    A = randmatrix(rng, N, k)
    d, o = 2 + 0.1j, -1.0
    ref = OB.project(A, lambda X: OO.tridiag_apply(X, d, o), w)
    plan = make_plan(backend)
    G = F.Tridiagonal1D(N, d, o, backend=backend)
    Q = F.PanelMatrix.from_array(A, plan=plan, w=w)
    p = F.npanels(Q)
    g = F.panelblas.project_group(Q) if group is None else group
    assert g == p or group is not None
    runs = -(-p // g)
    c0 = F._lib.counters()
    S = F.project(Q, G, group=group, schedule="fused")
    c1 = F._lib.counters()
    assert OB.isapprox(S, ref, OB.blastol(C16, N, k))
    assert c1["h2d_bytes"] - c0["h2d_bytes"] == (1 + runs) * N * k * 16
    assert np.array_equal(F.to_matrix(Q), A)
    assert all(pn.pins == 0 for pn in Q.panels)


def test_project_by_runs_over_a_disk_backed_basis(backend, tmp_path):
    rng = np.random.default_rng(960)
    N, k, w = 2000, 24, 4
    A = randmatrix(rng, N, k)
    d, o = 2 + 0.1j, -1.0
    ref = OB.project(A, lambda X: OO.tridiag_apply(X, d, o), w)
    slot = P.align_up(P.panel_bytes(N, w, np.dtype(C16)), P.SLAB_ALIGN)
    plan = F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=4 * slot,
                           scratch_dir=str(tmp_path), panel_width=w)
    G = F.Tridiagonal1D(N, d, o, backend=backend)
    Q = F.PanelMatrix.from_array(A, plan=plan, w=w)
    before = F.disk_reads(Q)
    for group in (2, None):
        S = F.project(Q, G, group=group, schedule="fused")
        assert OB.isapprox(S, ref, OB.blastol(C16, N, k))
    assert F.disk_reads(Q) > before
    assert np.array_equal(F.to_matrix(Q), A)


@pytest.mark.parametrize("w", (1, 4, 7, 23))
@pytest.mark.parametrize("nb", (1, 2))
def test_foreachpanel_closures_that_broadcast(backend, w, nb):
    """The closures of test/test_executor.jl:25-37 and test/gpu_smoke.jl:79-84
    on device buffers: panel .= 2 .* panel .+ T(j) (affine_), panel .*= 2,
    panel .= T(j) (fill_), and panel .+= / .-= another buffer, each exactly
    the oracle's, over a ragged last panel."""
    rng = np.random.default_rng(980 + w)
    N, k = 9, 23
    A, B = randmatrix(rng, N, k), randmatrix(rng, N, k)
    plan = make_plan(backend)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    F.foreachpanel(lambda j, panel: panel.affine_(2, j + 1), pm, nbuffers=nb)
    index = np.array([[c // w + 1 for c in range(k)]] * N, dtype=C16)
    expect = OB.affine(A, 2, 0) + index
    assert np.array_equal(F.to_matrix(pm), expect)

    def double(j, panel):
        panel *= 2
    F.foreachpanel(double, pm, nbuffers=nb)
    assert np.array_equal(F.to_matrix(pm), OB.scale(expect, 2.0))
    F.foreachpanel(lambda j, panel: panel.fill_(j + 1), pm, nbuffers=nb)
    assert np.array_equal(F.to_matrix(pm), index)
    X = F.PanelMatrix.from_array(B, plan=plan, w=w)

    def add(j, xb, yb):
        yb += xb
        yb -= 0.5
    sweep(add, readpanels(X), updatepanels(pm), nbuffers=nb)
    assert np.array_equal(F.to_matrix(pm), OB.affine(index + B, 1.0, -0.5))
    # A complex scalar: the schoolbook product, then the constant.
    a, b = 0.3 - 1.7j, 1.25 + 0.5j
    F.foreachpanel(lambda j, panel: panel.affine_(a, b), X, nbuffers=nb)
    assert np.array_equal(F.to_matrix(X), OB.affine(B, a, b))


def test_broadcast_closure_over_narrowed_storage(backend):
    # test/gpu_smoke.jl:115-122: the closure sees ComplexF64 buffers; the
    # results are narrowed to the ComplexF32 host tier on the way back.
    rng = np.random.default_rng(990)
    N, k, w = 300, 24, 4
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, host_eltype=np.complex64)
    Z = F.PanelMatrix.from_array(A, plan=plan, w=w)
    A32 = A.astype(np.complex64).astype(C16)
    assert np.array_equal(F.to_matrix(Z), A32)

    def double(j, panel):
        assert panel.dtype == np.dtype(C16)
        panel *= 2
    F.foreachpanel(double, Z)
    assert np.array_equal(F.to_matrix(Z), 2 * A32)


@pytest.mark.parametrize("schedule", ("reference", "fused"))
@pytest.mark.parametrize("resident", (False, True))
def test_no_device_allocation_after_warmup(backend, resident, schedule):
    """test/test_cuda.jl:100-111 on the library's own counter: after one run to
    fill the pools, a data sweep, a copying sweep, panelmul, gram, cholqr2 and
    project allocate no device memory at all (every workspace and staging
    buffer comes from the plan's pool, and no kernel stages scalars)."""
    rng = np.random.default_rng(1000)
    N, k, w = 4000, 48, 8
    A = randmatrix(rng, N, k)
    plan = make_plan(backend, device_budget=256 << 20, device_resident=resident, schedule=schedule)
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, backend=backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    Z = F.PanelMatrix.from_array(A, plan=plan, w=w)

    def double(j, panel):
        panel *= 2

    def copying(j, src, dst):
        backend.convert_copy(dst, src)
    steps = [lambda: sweep(double, updatepanels(X)),
             lambda: sweep(copying, readpanels(X), writepanels(Y)),
             lambda: F.panelmul(Y, G, X),
             lambda: F.gram(X),
             lambda: F.cholqr2(Z),
             lambda: F.project(Z, G)]
    for fn in steps:
        fn()
    backend.sync_all()
    for fn in steps:
        before = F._lib.counters()["device_allocs"]
        fn()
        backend.sync_all()
        assert F._lib.counters()["device_allocs"] == before
