"""The operator contract and check_operator, after test/test_operator.jl and
the test operators of test/operators.jl, on the CPU (jittered backend).

Besides the shipped DenseOperator, the operators here each keep or break one
part of the contract: one that only takes single columns, a matrix-free
circular convolution, one that reports a workspace, and five that break it
the ways check_operator has to catch (a transposed-not-adjoint adjoint, an
adjoint of the wrong shape, a panel result unlike the column results, a panel
claim that raises, a false Hermitian claim).
"""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import operator as O

from fakes import JitteryBackend, _np

C16 = np.dtype(np.complex128)


def randmatrix(rng, T, m, n):
    T = np.dtype(T)
    A = rng.standard_normal((m, n))
    if T.kind == "c":
        A = A + 1j * rng.standard_normal((m, n))
    return np.asfortranarray(A.astype(T))


def hermitian(A):
    return np.asfortranarray(np.triu(A) + np.triu(A, 1).conj().T + np.diag(np.diag(A).real - np.diag(A)))


BACKEND = JitteryBackend(jitter=1e-4)


def checkop(G, **kw):
    return F.check_operator(G, backend=BACKEND, **kw)


class _Dense:
    """Holds a host matrix on the device and applies it column by column."""

    def __init__(self, A):
        self.H = np.asfortranarray(A)
        self.A = BACKEND.todevice(self.H)
        self.shape = self.H.shape
        self.dtype = self.H.dtype

    def apply(self, Y, X):
        BACKEND.rightmul_gemm(Y, self.A, X)


class ColumnOperator(_Dense):
    """Dense, but takes one column at a time (not panel capable)."""

    def adjoint(self):
        return ColumnOperator(self.H.conj().T)


class Workspaced(_Dense):
    def __init__(self, A, nbytes):
        super().__init__(A)
        self.workspace_bytes = nbytes

    def adjoint(self):
        return Workspaced(self.H.conj().T, self.workspace_bytes)


class TransposedAdjoint(_Dense):
    def adjoint(self):
        return TransposedAdjoint(self.H.T)


class MisshapenAdjoint(_Dense):
    def adjoint(self):
        return MisshapenAdjoint(self.H.conj().T[:, :-1])


class WrongPanel(_Dense):
    """Claims whole panels, then gets the panel case wrong by a factor 2."""
    panel_capable = True

    def adjoint(self):
        return WrongPanel(self.H.conj().T)

    def apply(self, Y, X):
        BACKEND.rightmul_gemm(Y, self.A, X)
        if X.cols > 1:
            BACKEND.scale(Y, 2.0)


class UnhonouredPanel(_Dense):
    """Claims whole panels and then cannot take one."""
    panel_capable = True

    def adjoint(self):
        return UnhonouredPanel(self.H.conj().T)

    def apply(self, Y, X):
        if X.cols > 1:
            raise RuntimeError("UnhonouredPanel takes single columns only")
        BACKEND.rightmul_gemm(Y, self.A, X)


class FalselyHermitian(_Dense):
    ishermitian_op = True

    def adjoint(self):
        return ColumnOperator(self.H.conj().T)


class CircularConvolution:
    """Matrix free: (G x)_i = sum_j c[(i - j) mod N] x_j, its adjoint the
    conjugate reversed kernel."""

    def __init__(self, kernel, conjugated=False):
        self.kernel = np.asarray(kernel)
        self.conjugated = conjugated
        self.shape = (len(kernel), len(kernel))
        self.dtype = self.kernel.dtype

    def coefficient(self, i, j):
        N = len(self.kernel)
        return np.conj(self.kernel[(j - i) % N]) if self.conjugated else self.kernel[(i - j) % N]

    def dense(self):
        N = len(self.kernel)
        return np.array([[self.coefficient(i, j) for j in range(N)] for i in range(N)])

    def adjoint(self):
        return CircularConvolution(self.kernel, not self.conjugated)

    def apply(self, Y, X):
        def run():
            x, y = _np(X), _np(Y)
            N = len(self.kernel)
            for i in range(N):
                y[i, :] = sum(self.coefficient(i, j) * x[j, :] for j in range(N))
        BACKEND._launch(run)


class NarrowBackend(JitteryBackend):
    """A backend that cannot compute in complex128."""

    def supports_eltype(self, T):
        return np.dtype(T) != C16 and super().supports_eltype(T)


def contract_abiding_operators():
    rng = np.random.default_rng(1)
    dense = randmatrix(rng, C16, 12, 12)
    return [("dense", F.DenseOperator(dense, BACKEND)),
            ("hermitian", F.DenseOperator(hermitian(randmatrix(rng, C16, 12, 12)), BACKEND)),
            ("column", ColumnOperator(dense)),
            ("circular", CircularConvolution(randmatrix(rng, C16, 12, 1)[:, 0])),
            ("workspaced", Workspaced(dense, 4096))]


def test_trait_defaults():
    # test/test_operator.jl:3-16
    rng = np.random.default_rng(2)
    G = F.DenseOperator(randmatrix(rng, C16, 6, 6), BACKEND)
    assert F.workspace_bytes(G) == 0
    assert F.panel_capable(G)
    assert not F.ishermitian_op(G)
    assert F.ishermitian_op(F.DenseOperator(hermitian(randmatrix(rng, C16, 6, 6)), BACKEND))
    column = ColumnOperator(randmatrix(rng, C16, 6, 6))
    assert F.workspace_bytes(column) == 0
    assert not F.panel_capable(column)
    assert not F.ishermitian_op(column)
    assert F.workspace_bytes(Workspaced(randmatrix(rng, C16, 6, 6), 1 << 20)) == 1 << 20


@pytest.mark.parametrize("name,G", contract_abiding_operators(), ids=[n for n, _ in contract_abiding_operators()])
def test_check_operator_accepts(name, G):
    # test/test_operator.jl:28-32
    assert checkop(G)
    assert checkop(G, n=1)
    assert checkop(G.adjoint())


def test_check_operator_accepts_a_rectangular_operator():
    # test/test_operator.jl:34-36
    assert checkop(F.DenseOperator(randmatrix(np.random.default_rng(3), C16, 9, 5), BACKEND))


def test_check_operator_accepts_a_real_operator():
    # test/test_operator.jl:38-40
    assert checkop(F.DenseOperator(randmatrix(np.random.default_rng(4), np.float64, 8, 8), BACKEND))


def test_check_operator_catches_a_wrong_adjoint():
    # test/test_operator.jl:42-47: a real operator's transpose is its adjoint,
    # so that one has to pass.
    rng = np.random.default_rng(5)
    with pytest.raises(F.ArgumentError):
        checkop(TransposedAdjoint(randmatrix(rng, C16, 10, 10)))
    assert checkop(TransposedAdjoint(randmatrix(rng, np.float64, 10, 10)))


def test_check_operator_catches_a_wrong_shape():
    # test/test_operator.jl:49-53
    rng = np.random.default_rng(6)
    with pytest.raises(F.ArgumentError):
        checkop(MisshapenAdjoint(randmatrix(rng, C16, 10, 10)))
    with pytest.raises(F.ArgumentError):
        checkop(MisshapenAdjoint(randmatrix(rng, C16, 9, 5)))


def test_an_operator_has_to_match_both_row_counts():
    # test/test_operator.jl:55-62
    G = F.DenseOperator(randmatrix(np.random.default_rng(7), C16, 9, 5), BACKEND)
    assert O.assert_operator_shape(G, 9, 5) is None
    with pytest.raises(F.ArgumentError):
        O.assert_operator_shape(G, 5, 9)
    with pytest.raises(F.ArgumentError):
        O.assert_operator_shape(G, 9, 9)
    with pytest.raises(F.ArgumentError):
        O.assert_square_operator(G, 9)


def test_check_operator_catches_a_panel_and_column_mismatch():
    # test/test_operator.jl:64-66
    with pytest.raises(F.ArgumentError):
        checkop(WrongPanel(randmatrix(np.random.default_rng(8), C16, 10, 10)))


def test_check_operator_names_a_panel_claim_the_operator_cannot_honour():
    # test/test_operator.jl:68-73
    with pytest.raises(F.ArgumentError) as err:
        checkop(UnhonouredPanel(randmatrix(np.random.default_rng(9), C16, 10, 10)))
    assert "panel_capable" in str(err.value)
    assert "takes single columns only" in str(err.value)


def test_check_operator_catches_a_false_hermitian_claim():
    # test/test_operator.jl:75-80
    rng = np.random.default_rng(10)
    with pytest.raises(F.ArgumentError):
        checkop(FalselyHermitian(randmatrix(rng, C16, 10, 10)))
    assert checkop(FalselyHermitian(hermitian(randmatrix(rng, C16, 10, 10))))


def test_check_operator_rejects_nonsense_arguments():
    # test/test_operator.jl:82-86
    rng = np.random.default_rng(11)
    A = F.DenseOperator(randmatrix(rng, C16, 6, 6), BACKEND)
    with pytest.raises(F.ArgumentError):
        checkop(A, n=0)
    with pytest.raises(F.ArgumentError):
        F.check_operator(ColumnOperator(randmatrix(rng, C16, 6, 6)), backend=NarrowBackend())


def test_the_circular_convolution_operator_is_what_it_claims():
    # test/test_operator.jl:88-97
    rng = np.random.default_rng(12)
    G = CircularConvolution(randmatrix(rng, C16, 9, 1)[:, 0])
    dense = G.dense()
    xh = randmatrix(rng, C16, 9, 1)
    x = BACKEND.todevice(xh)
    y = BACKEND.alloc_device(C16, (9, 1))
    G.apply(y, x)
    assert np.allclose(BACKEND.to_host(y), dense @ xh, rtol=1e-13, atol=1e-13)
    assert np.allclose(G.adjoint().dense(), dense.conj().T)
