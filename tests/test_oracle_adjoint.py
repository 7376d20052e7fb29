"""Pins the oracle's adjoints with central finite differences (the QR
adjoint and project's, including the sensitivity to the operator's diagonal)."""
import numpy as np

from oracle import adjoint as OA
from oracle import operators as OO
from oracle import panelblas as OB


def cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_qr_pullback_matches_finite_differences():
    rng = np.random.default_rng(1)
    N, k = 60, 8
    Y = cplx(rng, N, k)
    Qb, Rb = cplx(rng, N, k), np.triu(cplx(rng, k, k))
    Q, R = OB.cholqr2(Y, 4)
    Ybar = OA.qr_pullback(Q, R, Qb, Rb)

    def loss(Y):
        Q, R = OB.cholqr2(Y, 4)
        return np.real(np.vdot(Qb, Q) + np.vdot(Rb, R))
    for _ in range(3):
        E = cplx(rng, N, k)
        h = 1e-6
        fd = (loss(Y + h * E) - loss(Y - h * E)) / (2 * h)
        assert abs(fd - np.real(np.vdot(Ybar, E))) <= 1e-6 * max(1.0, abs(fd))
    # The real-diagonal detail matters: dropping Re on diag(M) changes the result.
    M = R @ Rb.conj().T - Qb.conj().T @ Q
    L = np.tril(M, -1)
    wrong = (Qb + Q @ (L + L.conj().T + np.diag(np.diag(M)))) @ np.linalg.inv(R).conj().T
    assert not np.allclose(wrong, Ybar)


def test_project_pullback_matches_finite_differences():
    rng = np.random.default_rng(2)
    N, k = 50, 6
    Q = cplx(rng, N, k)
    d = 1 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    Sb = cplx(rng, k, k)

    def S_of(Q, d):
        return Q.conj().T @ OO.tridiag_apply(Q, 2 + 0.1j, -1.0, d)

    def loss(Q, d):
        return np.real(np.vdot(Sb, S_of(Q, d)))
    Qbar, dbar = OA.project_pullback(Q, lambda X: OO.tridiag_apply(X, 2 + 0.1j, -1.0, d),
                                     lambda X: OO.tridiag_apply(X, 2 + 0.1j, -1.0, d, adjoint=True), Sb)
    h = 1e-6
    E = cplx(rng, N, k)
    fd = (loss(Q + h * E, d) - loss(Q - h * E, d)) / (2 * h)
    assert abs(fd - np.real(np.vdot(Qbar, E))) <= 1e-6 * max(1.0, abs(fd))
    e = rng.standard_normal(N)
    fd = (loss(Q, d + h * e) - loss(Q, d - h * e)) / (2 * h)
    assert abs(fd - dbar @ e) <= 1e-6 * max(1.0, abs(fd))
