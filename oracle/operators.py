"""The benchmark configurations' operators, applied on the CPU.

TEST INFRASTRUCTURE (see oracle/__init__.py). Restated from the
configurations' own definitions: the periodic 1-D tridiagonal (diagonal d,
off-diagonals o, optional real diagonal parameters) and the periodic 3-D
7-point stencil with linear index i + n1 (j + n2 l).
"""
import numpy as np


def tridiag_apply(X, d, o, diag=None, adjoint=False):
    """(G x)_i = (d + diag_i) x_i + o (x_{i-1} + x_{i+1}), indices mod N; the
    adjoint conjugates the coefficients."""
    X = np.asarray(X)
    dd = np.full(X.shape[0], d, dtype=np.complex128)
    if diag is not None:
        dd = dd + np.asarray(diag, dtype=np.float64)
    oo = o
    if adjoint:
        dd = dd.conj()
        oo = np.conj(o)
    return dd[:, None] * X + oo * (np.roll(X, 1, axis=0) + np.roll(X, -1, axis=0))


def tridiag_dense(N, d, o, diag=None):
    """The same operator as an explicit matrix, entry by entry."""
    G = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        G[i, i] += d + (0.0 if diag is None else diag[i])
        G[i, (i - 1) % N] += o
        G[i, (i + 1) % N] += o
    return G


def helmholtz_apply(X, n, d, o=-1.0, adjoint=False):
    """7-point periodic stencil on n = (n1, n2, n3), column-major linear index."""
    n1, n2, n3 = n
    X = np.asarray(X)
    w = X.shape[1]
    V = X.reshape((n1, n2, n3, w), order="F")
    s = (np.roll(V, 1, 0) + np.roll(V, -1, 0) + np.roll(V, 1, 1) + np.roll(V, -1, 1) + np.roll(V, 1, 2)
         + np.roll(V, -1, 2))
    dd, oo = (np.conj(d), np.conj(o)) if adjoint else (d, o)
    Y = dd * V + oo * s
    return Y.reshape((n1 * n2 * n3, w), order="F")


def helmholtz_dense(n, d, o=-1.0):
    n1, n2, n3 = n
    N = n1 * n2 * n3
    G = np.zeros((N, N), dtype=np.complex128)
    for l in range(n3):
        for j in range(n2):
            for i in range(n1):
                p = i + n1 * (j + n2 * l)
                G[p, p] += d
                for (di, dj, dl) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    q = (i + di) % n1 + n1 * ((j + dj) % n2 + n2 * ((l + dl) % n3))
                    G[p, q] += o
    return G


def helmholtz_diagonal(k0=0.3, damping=0.05):
    """6 - k0^2 (1 + i damping): configurations 2 and 5."""
    return 6.0 - k0 * k0 * (1.0 + 1j * damping)
