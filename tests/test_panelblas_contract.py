"""The panel BLAS, after test/test_panelblas.jl, on two backends: the
jittered numpy backend with its jitter off (CPU, as the reference runs its
arithmetic sets on a quiet backend) and the HIP backend (GPU, marked gpu).
These are the testsets the GPU panel-BLAS tests do not already hold:
matrix-free, column-only and rectangular operators, the rightmul! forms,
the elementwise BLAS per element type, argument checks, the shifted
Cholesky's trigger, project's identities and the pool's balance.

Tolerances are the reference's: blastol(T, N, k) = 100 eps(real T) sqrt(N k)
in Julia's isapprox sense (test/setup.jl:170)."""
import numpy as np
import pytest

import funicular_hip as F

from fakes import JitteryBackend
from oracle import panelblas as OB
from test_operator import CircularConvolution

BLAS_WIDTHS = (23, 12, 7, 1)  # test/test_panelblas.jl:1
C16, C8, F64 = np.dtype(np.complex128), np.dtype(np.complex64), np.dtype(np.float64)
ELTYPES = (C16, C8, F64)  # test/setup.jl:130


@pytest.fixture(params=["jittered", pytest.param("hip", marks=pytest.mark.gpu)])
def be(request):
    if request.param == "hip":
        return F.hip_backend()
    return JitteryBackend(jitter=0.0)


def make_plan(backend, **kw):
    kw.setdefault("host_budget", 64 << 20)
    return F.ResidencyPlan(backend=backend, device_budget=256 << 20, **kw)


def randmatrix(rng, T, N, k):
    # test/setup.jl:148-149: uniform entries centred on zero
    T = np.dtype(T)
    A = rng.random((N, k)) - 0.5
    if T.kind == "c":
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(T))


def tol(T, N, k):
    return OB.blastol(np.dtype(T).type(0).real.dtype, N, k)


def close(x, y, rtol):
    return OB.isapprox(x, y, rtol)


class ColumnOperator:
    """Dense on the device, applied one column at a time (test/operators.jl:24-35)."""

    def __init__(self, dense):
        self.dense = dense
        self.shape = dense.shape
        self.dtype = dense.dtype

    def adjoint(self):
        return ColumnOperator(self.dense.adjoint())

    def apply(self, Y, X):
        self.dense.apply(Y, X)


def test_gram_with_narrowed_host_storage(be):
    # test/test_panelblas.jl:28-37
    rng = np.random.default_rng(1)
    N, k, w = 40, 23, 7
    A = randmatrix(rng, C16, N, k)
    X = F.PanelMatrix.from_array(A, plan=make_plan(be, host_eltype=C8), w=w)
    stored = A.astype(C8).astype(C16)
    assert close(F.gram(X), stored.conj().T @ stored, tol(C16, N, k))


def test_gram_rejects_mismatched_operands(be):
    # test/test_panelblas.jl:39-46
    plan = make_plan(be)
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=plan, w=5))
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=plan, w=7, dtype=C8))
    with pytest.raises(F.ArgumentError):
        F.gram(X, F.PanelMatrix(10, 20, plan=make_plan(be), w=7))


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", BLAS_WIDTHS)
@pytest.mark.parametrize("T", ELTYPES, ids=str)
def test_panelmul_against_dense(be, T, w, nb):
    # test/test_panelblas.jl:48-60
    rng = np.random.default_rng(2)
    N, k = 40, 23
    A = randmatrix(rng, T, N, k)
    reference = randmatrix(rng, T, N, N)
    dense = F.DenseOperator(reference, be)
    plan = make_plan(be)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w, dtype=T)
    F.panelmul(Y, dense, X, nbuffers=nb)
    assert close(F.to_matrix(Y), reference @ A, tol(T, N, k))
    F.panelmul(Y, ColumnOperator(dense), X, nbuffers=nb)
    assert close(F.to_matrix(Y), reference @ A, tol(T, N, k))


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_panelmul_with_a_matrix_free_operator(w):
    # test/test_panelblas.jl:62-73 (the circulant computes with numpy, so it
    # runs on the jittered backend)
    rng = np.random.default_rng(3)
    N, k = 40, 23
    A = randmatrix(rng, C16, N, k)
    G = CircularConvolution(randmatrix(rng, C16, N, 1)[:, 0])
    from test_operator import BACKEND
    plan = make_plan(BACKEND)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, X)
    assert close(F.to_matrix(Y), G.dense() @ A, tol(C16, N, k))


def test_panelmul_rejects_an_operator_that_does_not_fit(be):
    # test/test_panelblas.jl:75-83
    rng = np.random.default_rng(4)
    plan = make_plan(be)
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    Y = F.PanelMatrix(10, 20, plan=plan, w=7)
    with pytest.raises(F.ArgumentError):
        F.panelmul(Y, F.DenseOperator(randmatrix(rng, C16, 10, 9), be), X)
    with pytest.raises(F.ArgumentError):
        F.panelmul(Y, F.DenseOperator(randmatrix(rng, C16, 11, 11), be), X)
    with pytest.raises(F.ArgumentError):
        F.panelmul(X, F.DenseOperator(randmatrix(rng, C16, 10, 10), be), X)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", (23, 7, 1))
@pytest.mark.parametrize("mn", ((40, 25), (25, 40)), ids=str)
@pytest.mark.parametrize("T", ELTYPES, ids=str)
def test_panelmul_with_a_rectangular_operator(be, T, mn, w, nb):
    # test/test_panelblas.jl:85-103; the adjoint maps the other way, so the
    # two matrices swap roles
    m, n = mn
    rng = np.random.default_rng(5)
    k = 23
    A = randmatrix(rng, T, n, k)
    B = randmatrix(rng, T, m, k)
    reference = randmatrix(rng, T, m, n)
    dense = F.DenseOperator(reference, be)
    plan = make_plan(be)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix(m, k, plan=plan, w=w, dtype=T)
    t = tol(T, max(m, n), k)
    F.panelmul(Y, dense, X, nbuffers=nb)
    assert close(F.to_matrix(Y), reference @ A, t)
    F.panelmul(Y, ColumnOperator(dense), X, nbuffers=nb)
    assert close(F.to_matrix(Y), reference @ A, t)
    W = F.PanelMatrix.from_array(B, plan=plan, w=w)
    Z = F.PanelMatrix(n, k, plan=plan, w=w, dtype=T)
    F.panelmul(Z, dense.adjoint(), W, nbuffers=nb)
    assert close(F.to_matrix(Z), reference.conj().T @ B, t)


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_panelmul_applies_a_rectangular_operator_to_a_ghost(be, w):
    # test/test_panelblas.jl:105-115
    rng = np.random.default_rng(6)
    m, n, k = 40, 25, 23
    reference = randmatrix(rng, C16, m, n)
    plan = make_plan(be)
    omega = F.GhostPanels(C16, n, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y = F.PanelMatrix(m, k, plan=plan, w=w)
    F.panelmul(Y, F.DenseOperator(reference, be), omega)
    assert close(F.to_matrix(Y), reference @ F.to_matrix(omega), tol(C16, m, k))


def test_panelmul_rejects_a_rectangular_operator_that_does_not_fit(be):
    # test/test_panelblas.jl:117-128
    rng = np.random.default_rng(7)
    plan = make_plan(be)
    X = F.PanelMatrix(12, 20, plan=plan, w=7)
    Y = F.PanelMatrix(15, 20, plan=plan, w=7)
    G = F.DenseOperator(randmatrix(rng, C16, 15, 12), be)
    with pytest.raises(F.ArgumentError):
        F.panelmul(Y, F.DenseOperator(randmatrix(rng, C16, 12, 15), be), X)
    with pytest.raises(F.ArgumentError):
        F.panelmul(Y, G, F.PanelMatrix(12, 20, plan=plan, w=5))
    with pytest.raises(F.ArgumentError):
        F.panelmul(Y, G, F.PanelMatrix(12, 18, plan=plan, w=7))
    assert F.panelmul(Y, G, X) is Y


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", BLAS_WIDTHS)
@pytest.mark.parametrize("r", (12, 23, 31))
@pytest.mark.parametrize("T", ELTYPES, ids=str)
def test_rightmul_out_of_place_against_dense(be, T, r, w, nb):
    # test/test_panelblas.jl:130-142
    rng = np.random.default_rng(8)
    N, s = 40, 23
    A = randmatrix(rng, T, N, s)
    C = randmatrix(rng, T, s, r)
    plan = make_plan(be)
    src = F.PanelMatrix.from_array(A, plan=plan, w=w)
    dest = F.PanelMatrix(N, r, plan=plan, w=min(max(w - 3, 1), r), dtype=T)
    assert F.rightmul(dest, src, C, nbuffers=nb) is dest
    assert close(F.to_matrix(dest), A @ C, tol(T, N, max(s, r)))
    assert np.array_equal(F.to_matrix(src), A)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", BLAS_WIDTHS)
@pytest.mark.parametrize("T", ELTYPES, ids=str)
def test_rightmul_in_place_against_dense(be, T, w, nb):
    # test/test_panelblas.jl:144-158
    rng = np.random.default_rng(9)
    N, k = 40, 23
    A = randmatrix(rng, T, N, k)
    C = randmatrix(rng, T, k, k)
    plan = make_plan(be)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    assert F.rightmul(Y, C, nbuffers=nb) is Y
    assert close(F.to_matrix(Y), A @ C, tol(T, N, k))
    dest = F.PanelMatrix(N, k, plan=plan, w=w, dtype=T)
    F.rightmul(dest, F.PanelMatrix.from_array(A, plan=plan, w=w), C, nbuffers=nb)
    assert close(F.to_matrix(dest), F.to_matrix(Y), tol(T, N, k))


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_rightmul_is_the_same_to_the_bit_at_nbuffers_1_and_2(be, w):
    # test/test_panelblas.jl:160-180
    rng = np.random.default_rng(10)
    N, s, r = 40, 23, 12
    A, C, D = randmatrix(rng, C16, N, s), randmatrix(rng, C16, s, r), randmatrix(rng, C16, s, s)
    produced, inplace = [], []
    for nb in (1, 2):
        plan = make_plan(be)
        src = F.PanelMatrix.from_array(A, plan=plan, w=w)
        dest = F.PanelMatrix(N, r, plan=plan, w=min(max(w - 3, 1), r))
        produced.append(F.to_matrix(F.rightmul(dest, src, C, nbuffers=nb)))
        inplace.append(F.to_matrix(F.rightmul(F.PanelMatrix.from_array(A, plan=make_plan(be), w=w), D,
                                              nbuffers=nb)))
    assert produced[0].tobytes() == produced[1].tobytes()
    assert inplace[0].tobytes() == inplace[1].tobytes()


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_rightmul_takes_a_wrapped_factor(be, w):
    # test/test_panelblas.jl:182-194: a triangular factor, its inverse, and a
    # conjugate-transposed view (row-major in memory)
    rng = np.random.default_rng(11)
    N, k = 40, 23
    A = randmatrix(rng, C16, N, k)
    R = np.triu(randmatrix(rng, C16, k, k)) + 4 * np.eye(k)
    D = randmatrix(rng, C16, k, k).conj().T
    plan = make_plan(be)
    t = tol(C16, N, k)
    assert close(F.to_matrix(F.rightmul(F.PanelMatrix.from_array(A, plan=plan, w=w), R)), A @ R, t)
    Rinv = np.linalg.inv(R)
    assert close(F.to_matrix(F.rightmul(F.PanelMatrix.from_array(A, plan=plan, w=w), Rinv)), A @ Rinv, t)
    assert close(F.to_matrix(F.rightmul(F.PanelMatrix.from_array(A, plan=plan, w=w), D)), A @ D, t)


def test_rightmul_with_narrowed_host_storage(be):
    # test/test_panelblas.jl:196-211: both tiers hold S, so the product is
    # narrowed on the way back too, and the tolerance is the storage eltype's
    rng = np.random.default_rng(12)
    N, s, r, w = 40, 23, 12, 7
    A = randmatrix(rng, C16, N, s)
    C = randmatrix(rng, C16, s, r)
    plan = make_plan(be, host_eltype=C8)
    src = F.PanelMatrix.from_array(A, plan=plan, w=w)
    dest = F.PanelMatrix(N, r, plan=plan, w=5)
    F.rightmul(dest, src, C)
    stored = A.astype(C8).astype(C16)
    assert close(F.to_matrix(dest), stored @ C, tol(C8, N, s))


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_rightmul_reads_a_ghost_source(be, w):
    # test/test_panelblas.jl:213-224
    rng = np.random.default_rng(13)
    N, s, r = 40, 23, 12
    C = randmatrix(rng, C16, s, r)
    plan = make_plan(be)
    omega = F.GhostPanels(C16, N, s, plan=plan, seed=0xDEADBEEF, w=w)
    dest = F.PanelMatrix(N, r, plan=plan, w=min(max(w - 3, 1), r))
    reference = F.to_matrix(omega)
    assert close(F.to_matrix(F.rightmul(dest, omega, C)), reference @ C, tol(C16, N, s))
    assert np.array_equal(F.to_matrix(omega), reference)


def test_rightmul_rejects_factors_and_matrices_that_do_not_fit(be):
    # test/test_panelblas.jl:226-242
    rng = np.random.default_rng(14)
    N, s, r, w = 20, 12, 8, 5
    plan = make_plan(be)
    Y = F.PanelMatrix.from_array(randmatrix(rng, C16, N, s), plan=plan, w=w)
    dest = F.PanelMatrix(N, r, plan=plan, w=4)
    omega = F.GhostPanels(C16, N, s, plan=plan, w=w)
    for call in (lambda: F.rightmul(Y, randmatrix(rng, C16, s, r)),
                 lambda: F.rightmul(Y, randmatrix(rng, C16, r, s)),
                 lambda: F.rightmul(dest, Y, randmatrix(rng, C16, s, s)),
                 lambda: F.rightmul(Y, Y, randmatrix(rng, C16, s, s)),
                 lambda: F.rightmul(omega, randmatrix(rng, C16, s, s)),
                 lambda: F.rightmul(omega, Y, randmatrix(rng, C16, s, s)),
                 lambda: F.rightmul(dest, F.PanelMatrix(N + 1, s, plan=plan, w=w), randmatrix(rng, C16, s, r)),
                 lambda: F.rightmul(dest, F.PanelMatrix(N, s, plan=make_plan(be), w=w),
                                    randmatrix(rng, C16, s, r))):
        with pytest.raises(F.ArgumentError):
            call()


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", BLAS_WIDTHS)
@pytest.mark.parametrize("T", ELTYPES, ids=str)
def test_axpy_scale_dot_and_norm(be, T, w, nb):
    # test/test_panelblas.jl:244-260
    rng = np.random.default_rng(15)
    N, k = 40, 23
    A, B = randmatrix(rng, T, N, k), randmatrix(rng, T, N, k)
    alpha = (0.5 - 1.25j) if np.dtype(T).kind == "c" else 0.75
    plan = make_plan(be)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    t = tol(T, N, k)
    assert close(F.dot(X, Y, nbuffers=nb), np.vdot(A, B), t)
    assert close(F.dot(X, X, nbuffers=nb), np.vdot(A, A), t)
    assert close(F.norm(X, nbuffers=nb), np.linalg.norm(A), t)
    assert F.scale(Y, alpha, nbuffers=nb) is Y
    assert close(F.to_matrix(Y), alpha * B, t)
    assert F.axpy(alpha, X, Y, nbuffers=nb) is Y
    assert close(F.to_matrix(Y), alpha * A + alpha * B, t)


@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_reductions_are_the_same_to_the_bit_at_nbuffers_1_and_2(be, w):
    # test/test_panelblas.jl:262-271
    rng = np.random.default_rng(16)
    N, k = 40, 23
    plan = make_plan(be)
    X = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=w)
    Y = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=w)
    d1, d2 = F.dot(X, Y, nbuffers=1), F.dot(X, Y, nbuffers=2)
    assert np.array(d1).tobytes() == np.array(d2).tobytes()
    n1, n2 = F.norm(X, nbuffers=1), F.norm(X, nbuffers=2)
    assert np.array(n1).tobytes() == np.array(n2).tobytes()


def test_the_elementwise_operations_reject_mismatched_operands(be):
    # test/test_panelblas.jl:273-283
    plan = make_plan(be)
    X = F.PanelMatrix(10, 20, plan=plan, w=7)
    Y = F.PanelMatrix(10, 20, plan=plan, w=5)
    with pytest.raises(F.ArgumentError):
        F.axpy(2.0, X, Y)
    with pytest.raises(F.ArgumentError):
        F.axpy(2.0, X, X)
    with pytest.raises(F.ArgumentError):
        F.dot(X, Y)
    with pytest.raises(F.ArgumentError):
        F.norm(X, 1)


def test_the_shifted_fallback_fires_only_when_it_has_to():
    # test/test_panelblas.jl:327-345: a zero column puts an exact zero on the
    # diagonal, which LAPACK always refuses; -G cannot be shifted to health
    rng = np.random.default_rng(17)
    N, k = 40, 23
    healthy = randmatrix(rng, C16, N, k)
    G = healthy.conj().T @ healthy
    _, shifted = F.cholesky_upper(G, N, k)
    assert not shifted
    singular = G.copy()
    singular[:, -1] = 0
    singular[-1, :] = 0
    factor, shifted = F.cholesky_upper(singular, N, k)
    assert shifted
    assert np.array_equal(factor, np.triu(factor))
    with pytest.raises(F.ArgumentError):
        F.cholesky_upper(-G, N, k)


@pytest.mark.parametrize("nb", (1, 2))
@pytest.mark.parametrize("w", BLAS_WIDTHS)
def test_project_against_dense(be, w, nb):
    # test/test_panelblas.jl:347-360
    rng = np.random.default_rng(18)
    N, k = 40, 23
    reference = randmatrix(rng, C16, N, N)
    dense = F.DenseOperator(reference, be)
    A = randmatrix(rng, C16, N, k)
    Q = F.PanelMatrix.from_array(A, plan=make_plan(be), w=w)
    S = A.conj().T @ reference @ A
    assert close(F.project(Q, dense, nbuffers=nb), S, tol(C16, N, k))
    assert close(F.project(Q, ColumnOperator(dense), nbuffers=nb), S, tol(C16, N, k))
    assert np.array_equal(F.to_matrix(Q), A)


def test_project_matches_panelmul_followed_by_gram(be):
    # test/test_panelblas.jl:362-373
    rng = np.random.default_rng(19)
    N, k, w = 40, 23, 7
    dense = F.DenseOperator(randmatrix(rng, C16, N, N), be)
    plan = make_plan(be)
    Q = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=w)
    Z = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Z, dense, Q)
    assert close(F.project(Q, dense), F.gram(Q, Z), tol(C16, N, k))


def test_project_follows_the_operators_adjoint(be):
    # test/test_panelblas.jl:375-390
    rng = np.random.default_rng(20)
    N, k, w = 40, 23, 7
    reference = randmatrix(rng, C16, N, N)
    herm = reference + reference.conj().T
    A = randmatrix(rng, C16, N, k)
    Q = F.PanelMatrix.from_array(A, plan=make_plan(be), w=w)
    t = tol(C16, N, k)
    S = F.project(Q, F.DenseOperator(reference, be))
    assert close(F.project(Q, F.DenseOperator(np.asfortranarray(reference.conj().T), be)), S.conj().T, t)
    H = F.project(Q, F.DenseOperator(herm, be))
    assert close(H, H.conj().T, t)


def test_the_panel_blas_gives_its_buffers_back(be):
    # test/test_panelblas.jl:433-459
    rng = np.random.default_rng(21)
    N, k, w = 40, 23, 7
    plan = make_plan(be)
    X = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=w)
    Y = F.PanelMatrix.from_array(randmatrix(rng, C16, N, k), plan=plan, w=w)
    Z = F.PanelMatrix(N, 12, plan=plan, w=5)
    dense = F.DenseOperator(randmatrix(rng, C16, N, N), be)
    C, D = randmatrix(rng, C16, k, 12), randmatrix(rng, C16, k, k)

    def round_of_calls():
        F.gram(X, Y)
        F.norm(X)
        F.project(X, dense)
        F.rightmul(Z, X, C)
        F.rightmul(Y, D)
        F.cholqr2(Y)

    round_of_calls()
    before = F.device_bytes_allocated(plan)
    round_of_calls()
    assert F.device_bytes_allocated(plan) == before
