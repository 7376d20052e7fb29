"""The reference CPU backend's panel BLAS in its own step order, in FP64
numpy + LAPACK.

TEST INFRASTRUCTURE (see oracle/__init__.py). The reference CPU backend's
"device" is a plain Array and its kernels are LinearAlgebra calls
(src/kernels.jl:40, :73, :108; src/device_api.jl:291-299), so restating it
means the same products over the same row blocks in the same order.
"""
import numpy as np
import scipy.linalg

from . import plan as P


def gram(A, B=None, w=None):
    """src/panelblas.jl:33-59: sum over row blocks of height floor(N w / k) of
    A_b^H B_b, the first block overwriting (carry 0), later ones adding."""
    B = A if B is None else B
    N, k = A.shape
    w = k if w is None else w
    acc = None
    for (r0, r1) in P.row_blocks(N, P.row_block_height(N, k, w)):
        prod = A[r0:r1].conj().T @ B[r0:r1]
        acc = prod if acc is None else acc + prod
    return acc


def cholesky_upper(G, N, k):
    """src/panelblas.jl:148-155: upper factor of Hermitian(G) (upper triangle
    read), retried with G + s I, s = 11 N k eps ||G||_F. Returns (R, shifted);
    raises ValueError where the reference throws ArgumentError."""
    G = np.asarray(G)
    potrf, = scipy.linalg.get_lapack_funcs(("potrf",), (G,))
    c, info = potrf(G, lower=False, clean=True)
    if info == 0:
        return np.triu(c), False
    eps = np.finfo(G.dtype).eps
    shift = 11 * N * k * eps * np.linalg.norm(G)
    c, info = potrf(G + shift * np.eye(G.shape[0], dtype=G.dtype), lower=False, clean=True)
    if info != 0:
        raise ValueError("Gram matrix not positive definite even after the shift")
    return np.triu(c), True


def rdiv_upper(Y, R):
    """src/kernels.jl:108: Y R^-1 with R upper (right side, no transpose)."""
    # Y R^-1 = (R^-T Y^T)^T: solve R^T X^T = Y^T with R^T lower.
    return scipy.linalg.solve_triangular(R, Y.T, trans="T", lower=False).T


def rdiv_rows(Y, R, w=None):
    """src/panelblas.jl:157-169: the solve block by block over the row traversal."""
    N, k = Y.shape
    w = k if w is None else w
    out = np.empty_like(Y)
    for (r0, r1) in P.row_blocks(N, P.row_block_height(N, k, w)):
        out[r0:r1] = rdiv_upper(Y[r0:r1], R)
    return out


def cholqr_pass(Y, w=None):
    """src/panelblas.jl:141-146."""
    G = gram(Y, w=w)
    R, shifted = cholesky_upper(G, Y.shape[0], Y.shape[1])
    return rdiv_rows(Y, R, w), R, shifted


def cholqr2(Y, w=None):
    """src/panelblas.jl:130-139: returns (Q, R)."""
    Q, R, s1 = cholqr_pass(Y, w)
    Q, R2, s2 = cholqr_pass(Q, w)
    R = R2 @ R
    if s1 or s2:
        Q, R3, _ = cholqr_pass(Q, w)
        R = R3 @ R
    return Q, R


def rightmul(A, C):
    """src/kernels.jl:73 over row blocks: A C."""
    return A @ C


def project(Q, apply, w):
    """src/panelblas.jl:268-301: S = Q^H G Q in project's p (p + 1)-step
    schedule; `apply(X)` is G applied to a block of columns."""
    N, k = Q.shape
    ranges = P.panel_ranges(k, w)
    panels, labels = P.project_labels(len(ranges))
    S = np.full((k, k), np.nan, dtype=Q.dtype)
    image = None
    for panel, (applying, i, j) in zip(panels, labels):
        c0, c1 = ranges[panel]
        Qb = Q[:, c0:c1]
        if applying:
            image = apply(Qb)
        else:
            ri, rj = ranges[i], ranges[j]
            S[ri[0]:ri[1], rj[0]:rj[1]] = Qb.conj().T @ image
    return S


def panelmul(apply, X, w):
    """src/panelblas.jl:91-100: G applied panel by panel."""
    out = None
    for (c0, c1) in P.panel_ranges(X.shape[1], w):
        Yb = apply(X[:, c0:c1])
        if out is None:
            out = np.empty((Yb.shape[0], X.shape[1]), dtype=Yb.dtype)
        out[:, c0:c1] = Yb
    return out


def gram_panelwise(A, w):
    """X^H X from column panel pairs: block (i, j) of the upper triangle is
    A_i^H A_j, the lower triangle mirrored. Not in the reference (its gram is
    a row sweep); restated from the block identity (A^H A)_ij = A_i^H A_j."""
    N, k = A.shape
    G = np.zeros((k, k), dtype=A.dtype)
    ranges = P.panel_ranges(k, w)
    for j, (c0, c1) in enumerate(ranges):
        for (d0, d1) in ranges[:j + 1]:
            G[d0:d1, c0:c1] = A[:, d0:d1].conj().T @ A[:, c0:c1]
    return np.triu(G) + np.triu(G, 1).conj().T


def rdiv_panelwise(Y, R, w):
    """Y R^-1 by column panels, left looking: X_j = (Y_j - sum_{i<j} X_i R_ij)
    R_jj^-1. Not in the reference; restated from block back substitution."""
    X = np.array(Y, copy=True)
    ranges = P.panel_ranges(Y.shape[1], w)
    for j, (c0, c1) in enumerate(ranges):
        H = np.array(Y[:, c0:c1], copy=True)
        for (d0, d1) in ranges[:j]:
            H -= X[:, d0:d1] @ R[d0:d1, c0:c1]
        X[:, c0:c1] = rdiv_upper(H, R[c0:c1, c0:c1])
    return X


def cholqr2_panelwise(Y, w):
    """cholqr2 with gram_panelwise and rdiv_panelwise in place of the row
    sweeps (two passes, a third after a shift). Returns (Q, R)."""
    N, k = Y.shape
    Q = np.array(Y, copy=True)
    R_total = None
    shifted_any = False
    for npass in range(3):
        if npass == 2 and not shifted_any:
            break
        R, shifted = cholesky_upper(gram_panelwise(Q, w), N, k)
        shifted_any = shifted_any or shifted
        Q = rdiv_panelwise(Q, R, w)
        R_total = R if R_total is None else R @ R_total
    return Q, R_total


def scale(Y, alpha):
    """src/panelblas.jl:311-316 (rmul! on each panel): Julia's Complex * Real
    scales both parts, Complex * Complex is the schoolbook product, each part
    rounded once per product and sum. Spelled out on the real parts so that
    numpy's complex loops (which may fuse) play no part."""
    Y = np.asarray(Y)
    if not np.iscomplexobj(Y):
        return Y * Y.dtype.type(np.real(alpha))
    R = Y.real.dtype.type
    yr, yi = Y.real.copy(), Y.imag.copy()
    with np.errstate(invalid="ignore", over="ignore"):
        if not np.iscomplexobj(alpha):
            a = R(alpha)
            re, im = yr * a, yi * a
        else:
            ar, ai = R(alpha.real), R(alpha.imag)
            re = yr * ar - yi * ai
            im = yr * ai + yi * ar
    out = np.empty_like(Y)
    out.real, out.imag = re, im
    return out


def affine(Y, alpha, beta):
    """Julia's Y .= alpha .* Y .+ beta inside a foreachpanel closure
    (src/executor.jl:241-267, test/test_executor.jl:30-31): the product as
    scale forms it, then beta added to each part, one rounding per
    operation."""
    P = scale(Y, alpha)
    if not np.iscomplexobj(P):
        return P + P.dtype.type(np.real(beta))
    R = P.real.dtype.type
    b = complex(beta)
    out = np.empty_like(P)
    with np.errstate(invalid="ignore", over="ignore"):
        out.real, out.imag = P.real + R(b.real), P.imag + R(b.imag)
    return out


def axpy(alpha, X, Y):
    """src/panelblas.jl:323-331: Y + alpha X, panel by panel (the panels are
    independent, so the order does not matter)."""
    return np.asarray(Y) + np.asarray(alpha, dtype=np.asarray(X).dtype) * np.asarray(X)


def dot(A, B, w):
    """src/panelblas.jl:333-354: an N x w accumulator that the first panel
    overwrites with conj(A_j) .* B_j and later panels add to (columns
    1:width of each panel), then the sum of the accumulator."""
    N, k = A.shape
    acc = None
    for (c0, c1) in P.panel_ranges(k, w):
        prod = A[:, c0:c1].conj() * B[:, c0:c1]
        if acc is None:
            acc = prod.copy()
        else:
            acc[:, :c1 - c0] += prod
    return acc.sum() if acc is not None else A.dtype.type(0)


def norm(A, w):
    """src/panelblas.jl:356-370: the square root of the summed |a|^2
    accumulator, with no rescaling."""
    N, k = A.shape
    acc = None
    for (c0, c1) in P.panel_ranges(k, w):
        sq = np.abs(A[:, c0:c1]) ** 2
        if acc is None:
            acc = sq.copy()
        else:
            acc[:, :c1 - c0] += sq
    return np.sqrt(acc.sum()) if acc is not None else 0.0


def orthogonality_error(Q):
    """||Q^H Q - I||_F."""
    k = Q.shape[1]
    return np.linalg.norm(Q.conj().T @ Q - np.eye(k))


def blastol(dtype, N, k):
    """test/setup.jl:170: 100 eps sqrt(N k)."""
    return 100 * np.finfo(dtype).eps * np.sqrt(N * k)


def orthogonality_bound(dtype, N, k):
    """test/setup.jl:172: 100 N k eps."""
    return 100 * N * k * np.finfo(dtype).eps


def isapprox(x, y, rtol):
    """Julia's isapprox on arrays: ||x - y|| <= rtol max(||x||, ||y||)."""
    x = np.asarray(x)
    y = np.asarray(y)
    return np.linalg.norm(x - y) <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))
