"""scale / axpy / dot / norm (src/panelblas.jl:303-370) on the HIP backend
against the oracle; test/test_panelblas.jl:244-282 and
test/test_ghosts.jl:125-140 are the models."""
import numpy as np
import pytest

import funicular_hip as F
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

WIDTHS = (23, 12, 7, 1)
EL = (np.complex128, np.complex64, np.float64, np.float32)


def randmatrix(rng, N, k, dtype):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(dtype, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(dtype))


def make_plan(backend, **kw):
    kw.setdefault("device_budget", 256 * 2 ** 20)
    kw.setdefault("host_budget", 256 * 2 ** 20)
    return F.ResidencyPlan(backend=backend, **kw)


def alpha_of(T):
    return T(0.5 - 1.25j) if np.issubdtype(T, np.complexfloating) else T(0.75)


@pytest.mark.parametrize("T", EL)
@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("nb", (1, 2))
def test_elementwise_against_the_oracle(backend, T, w, nb):
    print(f"elementwise T={np.dtype(T).name} w={w} nb={nb}", flush=True)
    rng = np.random.default_rng(100 * w + nb)
    N, k = 40, 23
    A, B = randmatrix(rng, N, k, T), randmatrix(rng, N, k, T)
    alpha = alpha_of(T)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    tol = OB.blastol(T, N, k)
    d = F.dot(X, Y, nbuffers=nb)
    assert isinstance(d, np.dtype(T).type)
    assert abs(d - OB.dot(A, B, w)) <= tol * abs(OB.dot(A, B, w))
    dd = F.dot(X, X, nbuffers=nb)
    assert abs(dd - OB.dot(A, A, w)) <= tol * abs(OB.dot(A, A, w))
    n = F.norm(X, nbuffers=nb)
    assert isinstance(n, np.finfo(T).dtype.type)
    assert abs(n - OB.norm(A, w)) <= tol * OB.norm(A, w)
    assert F.scale(Y, alpha, nbuffers=nb) is Y
    # The schoolbook product with one rounding per product and sum: bit-exact.
    assert np.array_equal(F.to_matrix(Y), OB.scale(B, alpha))
    assert F.axpy(alpha, X, Y, nbuffers=nb) is Y
    assert OB.isapprox(F.to_matrix(Y), OB.axpy(alpha, A, OB.scale(B, alpha)), tol)
    assert np.array_equal(F.to_matrix(X), A)


def test_scale_by_a_real_scalar_keeps_infinities(backend):
    # Complex * Real scales each part; Complex * Complex forms inf * 0.
    plan = make_plan(backend)
    A = np.asfortranarray(np.array([[complex(np.inf, 0.0), 1 - 2j], [3j, -0.0 + 0j]]))
    X = F.PanelMatrix.from_array(A.copy(), plan=plan, w=1)
    F.scale(X, 2.0)
    assert np.array_equal(F.to_matrix(X), OB.scale(A, 2.0))
    Y = F.PanelMatrix.from_array(A.copy(), plan=plan, w=2)
    F.scale(Y, 2.0 + 0j)
    got, want = F.to_matrix(Y), OB.scale(A, 2.0 + 0j)
    assert np.isnan(got[0, 0].imag) and np.isnan(want[0, 0].imag)
    assert np.array_equal(np.nan_to_num(got), np.nan_to_num(want))


@pytest.mark.parametrize("w", WIDTHS)
def test_reductions_same_bits_at_nbuffers_1_and_2(backend, w):
    rng = np.random.default_rng(w)
    A, B = randmatrix(rng, 40, 23, np.complex128), randmatrix(rng, 40, 23, np.complex128)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    assert F.dot(X, Y, nbuffers=1) == F.dot(X, Y, nbuffers=2)
    assert F.norm(X, nbuffers=1) == F.norm(X, nbuffers=2)


@pytest.mark.parametrize("N", (4095, 4096, 4097, 100_003))
def test_reductions_over_many_row_chunks(backend, N):
    # Row counts either side of one workgroup's chunk, and many chunks.
    print(f"reductions N={N}", flush=True)
    rng = np.random.default_rng(N)
    k, w = 9, 4
    A, B = randmatrix(rng, N, k, np.complex128), randmatrix(rng, N, k, np.complex128)
    plan = make_plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Y = F.PanelMatrix.from_array(B, plan=plan, w=w)
    tol = OB.blastol(np.complex128, N, k)
    ref = OB.dot(A, B, w)
    assert abs(F.dot(X, Y) - ref) <= tol * abs(ref)
    assert abs(F.norm(Y) - OB.norm(B, w)) <= tol * OB.norm(B, w)
    assert F.dot(X, Y) == F.dot(X, Y)  # deterministic


def test_elementwise_rejects_mismatched_operands(backend):
    plan = make_plan(backend)
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
    R = F.PanelMatrix(10, 20, plan=plan, w=7, dtype=np.float64)
    with pytest.raises(F.ArgumentError):
        F.scale(R, 1j)


def test_ghosts_refuse_scale_and_axpy(backend):
    plan = make_plan(backend)
    N, k, w = 16, 6, 4
    O = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    with pytest.raises(F.ArgumentError):
        F.scale(O, 2)
    with pytest.raises(F.ArgumentError):
        F.axpy(1.0, Y, O)
    assert all(p.pins == 0 for p in O.panels)
    # Reading a ghost is fine: its norm is that of the materialized matrix.
    A = F.to_matrix(O)
    assert abs(F.norm(O) - OB.norm(A, w)) <= OB.blastol(np.complex128, N, k) * OB.norm(A, w)


def test_reductions_at_a_large_ghost_basis(backend):
    # Size-independent identities at N = 2^22: norm^2 = dot(X, X),
    # dot(X, a X) = a |X|^2, and norm(scale(X, a)) = |a| norm(X).
    N, k, w = 1 << 22, 32, 16
    print(f"large reductions N={N} k={k}", flush=True)
    plan = F.ResidencyPlan(backend=backend, device_budget=16 * 2 ** 30, host_budget=16 * 2 ** 30)
    O = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    X = F.PanelMatrix(N, k, plan=plan, w=w)
    F.copyto(X, O)
    n = F.norm(X)
    dd = F.dot(X, X)
    assert abs(n * n - dd.real) <= 1e-13 * dd.real and dd.imag == 0.0
    # Entries are standard complex normals: |X|^2 is close to N k.
    assert abs(dd.real / (N * k) - 1.0) < 0.01
    a = 0.6 - 0.8j
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.copyto(Y, O)
    F.scale(Y, a)
    assert abs(F.dot(X, Y) - a * dd.real) <= 1e-12 * dd.real
    assert abs(F.norm(Y) - abs(a) * n) <= 1e-12 * n
    F.axpy(-a, X, Y)
    assert F.norm(Y) <= 1e-14 * n
    for m in (X, Y):
        F.free(m)
