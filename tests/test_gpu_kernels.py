"""The HIP kernels against the oracle, through the C ABI (src/kernels.jl
semantics; test/test_kernels.jl is the model)."""
import numpy as np
import scipy.linalg
import pytest

from oracle import ghost as OG
from oracle import operators as OO
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

EL = [np.complex128, np.complex64, np.float64, np.float32]


def randmatrix(rng, N, k, dtype):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(dtype, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(dtype))


def tol(dtype, N, k):
    return OB.blastol(dtype, N, k)


@pytest.mark.parametrize("T", EL)
def test_gram_accumulate(backend, T):
    rng = np.random.default_rng(11)
    A, B = randmatrix(rng, 12, 5, T), randmatrix(rng, 12, 4, T)
    Ad, Bd = backend.todevice(A), backend.todevice(B)
    C = backend.alloc_device(T, (5, 4))
    backend.gram_accumulate(C, Ad, Bd, 1.0, 0.0)
    assert OB.isapprox(backend.to_host(C), A.conj().T @ B, tol(T, 12, 5))
    backend.gram_accumulate(C, Ad, Bd, 1.0, 1.0)
    assert OB.isapprox(backend.to_host(C), 2 * (A.conj().T @ B), tol(T, 12, 5))
    # A zero beta overwrites whatever the buffer held, NaN included.
    backend.fill(C, np.nan)
    backend.gram_accumulate(C, Ad, Bd, 1.0, 0.0)
    assert OB.isapprox(backend.to_host(C), A.conj().T @ B, tol(T, 12, 5))
    # Views: a row block and a sub-block of the accumulator.
    Cf = backend.alloc_device(T, (5, 4))
    backend.fill(Cf, 0)
    backend.gram_accumulate(Cf.view((0, 3), (0, 2)), Ad.view((0, 6), (0, 3)), Bd.view((0, 6), (0, 2)), 1.0, 0.0)
    H = backend.to_host(Cf)
    assert OB.isapprox(H[:3, :2], A[:6, :3].conj().T @ B[:6, :2], tol(T, 6, 3))
    assert np.all(H[3:, :] == 0) and np.all(H[:, 2:] == 0)


@pytest.mark.parametrize("T", [np.complex128, np.float64])
@pytest.mark.parametrize("m,k", [(1, 1), (40, 23), (1000, 64), (5000, 100), (3001, 200), (257, 33)])
def test_gram_shapes_and_hermitian_exactness(backend, T, m, k):
    rng = np.random.default_rng(m * 1000 + k)
    A = randmatrix(rng, m, k, T)
    Ad = backend.todevice(A)
    C = backend.alloc_device(T, (k, k))
    backend.gram_accumulate(C, Ad, Ad, 1.0, 0.0)
    H = backend.to_host(C)
    assert OB.isapprox(H, A.conj().T @ A, tol(T, m, k))
    assert np.array_equal(H, H.conj().T)
    # A different B (general tiles) on the same data agrees to tolerance.
    Bd = backend.todevice(A.copy(order="F"))
    C2 = backend.alloc_device(T, (k, k))
    backend.gram_accumulate(C2, Ad, Bd, 1.0, 0.0)
    assert OB.isapprox(backend.to_host(C2), H, tol(T, m, k))


def test_gram_complex_alpha_beta(backend):
    rng = np.random.default_rng(12)
    T = np.complex128
    A, B = randmatrix(rng, 300, 70, T), randmatrix(rng, 300, 70, T)
    C0 = randmatrix(rng, 70, 70, T)
    Ad, Bd, Cd = backend.todevice(A), backend.todevice(B), backend.todevice(C0)
    al, be = 0.5 - 1.25j, -0.75 + 0.5j
    backend.gram_accumulate(Cd, Ad, Bd, al, be)
    assert OB.isapprox(backend.to_host(Cd), al * (A.conj().T @ B) + be * C0, tol(T, 300, 70))
    # Same operand, complex alpha and beta: the lower triangle is still right.
    Cd = backend.todevice(C0)
    backend.gram_accumulate(Cd, Ad, Ad, al, be)
    assert OB.isapprox(backend.to_host(Cd), al * (A.conj().T @ A) + be * C0, tol(T, 300, 70))


def test_gram_is_deterministic(backend):
    rng = np.random.default_rng(13)
    A = randmatrix(rng, 20000, 96, np.complex128)
    Ad = backend.todevice(A)
    outs = []
    for _ in range(3):
        C = backend.alloc_device(np.complex128, (96, 96))
        backend.gram_accumulate(C, Ad, Ad, 1.0, 0.0)
        outs.append(backend.to_host(C))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


def test_gram_mixed_precision(backend):
    rng = np.random.default_rng(14)
    A = randmatrix(rng, 4096, 80, np.complex128)
    Ad = backend.todevice(A)
    C = backend.alloc_device(np.complex128, (80, 80))
    backend.gram_accumulate(C, Ad, Ad, 1.0, 0.0, compute=np.complex64)
    ref = A.conj().T @ A
    err = np.linalg.norm(backend.to_host(C) - ref) / np.linalg.norm(ref)
    assert err < 1e-5
    assert err > 1e-12  # it really ran in FP32


@pytest.mark.parametrize("compute", [None, np.complex64])
@pytest.mark.parametrize("m,ka,kb", [(3000, 128, 64), (2500, 130, 70), (700, 64, 200)])
def test_gram_cross_views(backend, m, ka, kb, compute):
    """A^H B for two distinct ComplexF64 blocks held as views of taller
    matrices (leading dimensions 1037 and 1500 rows past m), with alpha and
    beta, in FP64 and in FP32 products: full 64-column tiles, whose full row
    slices take the unguarded fetch, beside ragged tiles and partial last
    slices, which take the guarded one."""
    rng = np.random.default_rng(m + ka + kb)
    Abig = randmatrix(rng, m + 1037, ka + 3, np.complex128)
    Bbig = randmatrix(rng, m + 1500, kb + 1, np.complex128)
    C0 = randmatrix(rng, ka, kb, np.complex128)
    Ad, Bd = backend.todevice(Abig), backend.todevice(Bbig)
    A = Abig[5:5 + m, 2:2 + ka]
    B = Bbig[9:9 + m, 1:1 + kb]
    Cd = backend.todevice(C0)
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    backend.gram_accumulate(Cd, Ad.view((5, 5 + m), (2, 2 + ka)), Bd.view((9, 9 + m), (1, 1 + kb)),
                            alpha, beta, compute=compute)
    ref = alpha * (A.conj().T @ B) + beta * C0
    err = np.linalg.norm(backend.to_host(Cd) - ref) / np.linalg.norm(ref)
    if compute is None:
        assert err < 1e-14
    else:
        assert 1e-12 < err < 1e-5


@pytest.mark.parametrize("compute", [None, np.complex64])
@pytest.mark.parametrize("m,k", [(3001, 130), (2048, 64), (5000, 200)])
def test_gram_hermitian_row_block_view(backend, m, k, compute):
    """A^H A of a row block held as a view of a taller, wider matrix (the
    form every row sweep hands the Gram), in FP64 and FP32 products: the
    diagonal tiles, which stage one operand, and the partial last slices of
    the row chunks. The result is exactly Hermitian."""
    rng = np.random.default_rng(m + k)
    big = randmatrix(rng, m + 999, k + 2, np.complex128)
    A = big[7:7 + m, 1:1 + k]
    Ad = backend.todevice(big).view((7, 7 + m), (1, 1 + k))
    C = backend.alloc_device(np.complex128, (k, k))
    backend.gram_accumulate(C, Ad, Ad, 1.0, 0.0, compute=compute)
    H = backend.to_host(C)
    ref = A.conj().T @ A
    err = np.linalg.norm(H - ref) / np.linalg.norm(ref)
    assert np.array_equal(H, H.conj().T)
    assert err < (1e-14 if compute is None else 1e-5)


@pytest.mark.parametrize("T", EL)
def test_rightmul_gemm(backend, T):
    rng = np.random.default_rng(15)
    A, C = randmatrix(rng, 12, 5, T), randmatrix(rng, 5, 4, T)
    Ad, Cd = backend.todevice(A), backend.todevice(C)
    dst = backend.alloc_device(T, (12, 4))
    backend.fill(dst, np.nan)
    backend.rightmul_gemm(dst, Ad, Cd)
    assert OB.isapprox(backend.to_host(dst), A @ C, tol(T, 12, 5))
    backend.fill(dst, 0)
    backend.rightmul_gemm(dst.view((0, 7), None), Ad.view((0, 7), None), Cd)
    H = backend.to_host(dst)
    assert OB.isapprox(H[:7], A[:7] @ C, tol(T, 7, 5))
    assert np.all(H[7:] == 0)


@pytest.mark.parametrize("m,s,r", [(1000, 256, 256), (777, 70, 33), (64, 1, 1)])
def test_rightmul_gemm_shapes(backend, m, s, r):
    rng = np.random.default_rng(m + s + r)
    A, C = randmatrix(rng, m, s, np.complex128), randmatrix(rng, s, r, np.complex128)
    dst = backend.alloc_device(np.complex128, (m, r))
    backend.rightmul_gemm(dst, backend.todevice(A), backend.todevice(C))
    assert OB.isapprox(backend.to_host(dst), A @ C, tol(np.complex128, m, s))


@pytest.mark.parametrize("T", EL)
@pytest.mark.parametrize("k", (5, 64, 130))
def test_rdiv_upper_folded_reuses_the_factor(backend, T, k):
    # A first solve folds R into the workspace; later blocks reuse it (R^H too).
    rng = np.random.default_rng(k)
    R = np.asfortranarray(np.triu(randmatrix(rng, k, k, T)) + 4 * np.eye(k, dtype=T))
    Rd = backend.todevice(R)
    work = backend.alloc_device(np.uint8, (backend.trsm_workspace_bytes(k, T), 1))
    for adjoint in (False, True):
        Rs = R.conj().T if adjoint else R
        blocks = [randmatrix(rng, m, k, T) for m in (300, 77, 1)]
        dev = [backend.todevice(Y) for Y in blocks]
        if adjoint:
            backend.rdiv_upper_adjoint(dev[0], Rd, work)
        else:
            backend.rdiv_upper(dev[0], Rd, work)
        for d in dev[1:]:
            backend.rdiv_upper_folded(d, work, adjoint=adjoint)
        for Y, d in zip(blocks, dev):
            X = backend.to_host(d)
            assert np.linalg.norm(X @ Rs - Y) <= tol(T, Y.shape[0], k) * np.linalg.norm(Y)
    if T == np.complex128:
        Y = randmatrix(rng, 200, k, T)
        d0, d1 = backend.todevice(Y), backend.todevice(Y)
        w32 = backend.alloc_device(np.uint8, (backend.trsm_workspace_bytes(k, np.complex64), 1))
        backend.rdiv_upper(d0, Rd, w32, compute=np.complex64)
        backend.rdiv_upper_folded(d1, w32, compute=np.complex64)
        assert np.array_equal(backend.to_host(d0), backend.to_host(d1))


def hpd(rng, k, T, cond=1e3):
    Qm = np.linalg.qr(randmatrix(rng, k, k, T))[0]
    G = (Qm * np.logspace(0, -np.log10(cond), k)) @ Qm.conj().T
    return np.asfortranarray(((G + G.conj().T) / 2).astype(T))


@pytest.mark.parametrize("T", EL)
@pytest.mark.parametrize("k", (1, 5, 31, 32, 33, 64, 100, 512))
def test_potrf_upper_against_lapack(backend, T, k):
    # The lower triangle of G holds NaN: only the upper one may be read.
    rng = np.random.default_rng(k)
    G = hpd(rng, k, T)
    Gn = G.copy()
    Gn[np.tril_indices(k, -1)] = np.nan
    Gd = backend.todevice(Gn)
    Rd = backend.alloc_device(T, (k, k))
    backend.fill(Rd, np.nan)
    info = backend.alloc_device(np.int32, (1, 1))
    backend.potrf_upper(Rd, Gd, info)
    assert int(backend.to_host(info)[0, 0]) == 0
    R = backend.to_host(Rd)
    ref = scipy.linalg.cholesky(G.astype(np.complex128 if np.iscomplexobj(G) else np.float64), lower=False)
    assert np.all(np.tril(R, -1) == 0)
    assert np.all(np.diag(R).real > 0)
    eps = np.finfo(T).eps
    assert np.linalg.norm(R - ref) <= 50 * eps * k * np.linalg.norm(ref) * 1e3 ** 0.5
    assert np.linalg.norm(R.conj().T @ R - G) <= 50 * eps * k * np.linalg.norm(G)


def test_potrf_upper_narrows_and_reports_the_first_bad_pivot(backend):
    rng = np.random.default_rng(3)
    k = 100
    G = hpd(rng, k, np.complex128)
    # FP64 Gram factored in FP32, as the mixed first pass does.
    Gd = backend.todevice(G)
    R32 = backend.alloc_device(np.complex64, (k, k))
    info = backend.alloc_device(np.int32, (1, 1))
    backend.potrf_upper(R32, Gd, info)
    assert int(backend.to_host(info)[0, 0]) == 0
    R = backend.to_host(R32).astype(np.complex128)
    assert np.linalg.norm(R.conj().T @ R - G) <= 1e-5 * np.linalg.norm(G)
    # An indefinite leading minor of order 38 (1-based): LAPACK's info.
    B = G.copy()
    B[37, 37] = -1.0
    potrf, = scipy.linalg.get_lapack_funcs(("potrf",), (B,))
    _, lapack_info = potrf(B, lower=False)
    Rd = backend.alloc_device(np.complex128, (k, k))
    backend.potrf_upper(Rd, backend.todevice(B), info)
    assert int(backend.to_host(info)[0, 0]) == lapack_info == 38
    # A NaN on the diagonal stops it there too; a clean matrix resets info.
    Cn = G.copy()
    Cn[70, 70] = np.nan
    backend.potrf_upper(Rd, backend.todevice(Cn), info)
    assert int(backend.to_host(info)[0, 0]) == 71
    backend.potrf_upper(Rd, Gd, info)
    assert int(backend.to_host(info)[0, 0]) == 0


@pytest.mark.parametrize("T", EL)
def test_rdiv_upper(backend, T):
    rng = np.random.default_rng(16)
    Y = randmatrix(rng, 7, 5, T)
    R = np.asfortranarray(np.triu(randmatrix(rng, 5, 5, T)) + 4 * np.eye(5, dtype=T))
    Rd = backend.todevice(R)
    Yd = backend.todevice(Y)
    backend.rdiv_upper(Yd, Rd)
    assert OB.isapprox(backend.to_host(Yd), OB.rdiv_upper(Y, R), tol(T, 7, 5))
    # A row-strided view: rows outside it are untouched.
    S = backend.todevice(Y)
    backend.rdiv_upper(S.view((0, 3), None), Rd)
    H = backend.to_host(S)
    assert OB.isapprox(H[:3], OB.rdiv_upper(Y[:3], R), tol(T, 3, 5))
    assert np.array_equal(H[3:], Y[3:])


@pytest.mark.parametrize("compute", [None, np.complex64])
def test_rdiv_upper_bottom_row_block_of_a_taller_matrix(backend, compute):
    """Row blocks of a 4096 x 128 ComplexF64 matrix (an allocation of exactly
    8 MiB, so it ends on a 2 MiB page boundary), solved as views with
    ld = 4096: the bottom block of 1000 rows ends its last workgroup 24 rows
    past the matrix, and those rows must not be read past the allocation.
    Each block's result matches the solve of a standalone copy of it, and the
    rows outside the view are untouched."""
    rng = np.random.default_rng(31)
    N, k = 4096, 128
    Y = randmatrix(rng, N, k, np.complex128)
    R = np.asfortranarray(np.triu(randmatrix(rng, k, k, np.complex128)) + 8 * np.eye(k))
    Rd = backend.todevice(R)
    for r0, r1 in ((N - 1000, N), (0, 1000), (1000, 3001)):
        Yd = backend.todevice(Y)
        backend.rdiv_upper(Yd.view((r0, r1), None), Rd, compute=compute)
        alone = backend.todevice(np.asfortranarray(Y[r0:r1]))
        backend.rdiv_upper(alone, Rd, compute=compute)
        H = backend.to_host(Yd)
        assert np.array_equal(H[r0:r1], backend.to_host(alone)), (r0, r1)
        assert np.array_equal(H[:r0], Y[:r0]) and np.array_equal(H[r1:], Y[r1:])
        bar = 1e-5 if compute is not None else tol(np.complex128, r1 - r0, k)
        assert np.linalg.norm(H[r0:r1] @ R - Y[r0:r1]) <= bar * np.linalg.norm(Y[r0:r1])


@pytest.mark.parametrize("m,k", [(1000, 64), (5000, 512), (333, 45), (100, 1), (200, 96)])
def test_rdiv_upper_shapes(backend, m, k):
    rng = np.random.default_rng(m + k)
    T = np.complex128
    Y = randmatrix(rng, m, k, T)
    G = Y.conj().T @ Y
    R = np.asfortranarray(np.linalg.cholesky(G).conj().T)
    Yd = backend.todevice(Y)
    backend.rdiv_upper(Yd, backend.todevice(R))
    X = backend.to_host(Yd)
    ref = OB.rdiv_upper(Y, R)
    assert OB.isapprox(X, ref, 1e-12)
    assert np.linalg.norm(X @ R - Y) <= 1e-13 * np.linalg.norm(Y) * np.sqrt(k)


def test_rdiv_upper_mixed(backend):
    rng = np.random.default_rng(17)
    Y = randmatrix(rng, 2000, 100, np.complex128)
    R = np.asfortranarray(np.linalg.cholesky(Y.conj().T @ Y).conj().T)
    Yd = backend.todevice(Y)
    backend.rdiv_upper(Yd, backend.todevice(R), compute=np.complex64)
    X = backend.to_host(Yd)
    err = np.linalg.norm(X - OB.rdiv_upper(Y, R)) / np.linalg.norm(X)
    assert 1e-12 < err < 1e-4


@pytest.mark.parametrize("T", EL)
def test_convert_copy_and_fill(backend, T):
    rng = np.random.default_rng(18)
    A = randmatrix(rng, 9, 5, T)
    src = backend.todevice(A)
    dst = backend.alloc_device(T, (9, 5))
    backend.convert_copy(dst, src)
    assert np.array_equal(backend.to_host(dst), A)
    S = np.complex64 if np.issubdtype(T, np.complexfloating) else np.float32
    narrow = backend.alloc_device(S, (9, 5))
    backend.convert_copy(narrow, src)
    assert np.array_equal(backend.to_host(narrow), A.astype(S))
    backend.fill(dst, 0)
    backend.convert_copy(dst.view(None, (0, 2)), src.view(None, (3, 5)))
    backend.convert_copy(dst.view((0, 3), (2, 5)), src.view((6, 9), (0, 3)))
    H = backend.to_host(dst)
    assert np.array_equal(H[:, :2], A[:, 3:5])
    assert np.array_equal(H[:3, 2:5], A[6:9, :3])


@pytest.mark.parametrize("T", EL)
def test_ghost_fill_is_bit_exact(backend, T):
    N, k, seed = 5000, 6, 0xDEADBEEF
    d = backend.alloc_device(T, (N, k))
    backend.ghost_fill(d, 0, 1, seed)
    H = backend.to_host(d)
    ref = OG.ghost(T, 0, N, 1, k, seed)
    assert np.array_equal(H, ref)
    # A row block and a column offset regenerate the same entries.
    b = backend.alloc_device(T, (37, 2))
    backend.ghost_fill(b, 1234, 4, seed)
    assert np.array_equal(backend.to_host(b), ref[1234:1271, 3:5])


def test_ghost_fill_large_bit_exact(backend):
    # Enough draws that tail and wedge paths are taken many times.
    N = 400000
    d = backend.alloc_device(np.complex128, (N, 2))
    backend.ghost_fill(d, 0, 7, 12345)
    H = backend.to_host(d)
    assert np.array_equal(H, OG.ghost(np.complex128, 0, N, 7, 2, 12345))
    assert np.abs(H.real).max() * np.sqrt(2) > 3.7


def test_stencils(backend):
    import funicular_hip as F
    rng = np.random.default_rng(19)
    N = 1000
    X = randmatrix(rng, N, 3, np.complex128)
    diag = 1 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    Xd = backend.todevice(X)
    Yd = backend.alloc_device(np.complex128, (N, 3))
    for dg in (None, diag):
        G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, diag=dg, backend=backend)
        G.apply(Yd, Xd)
        assert np.allclose(backend.to_host(Yd), OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg), rtol=1e-14, atol=1e-14)
        G.adjoint().apply(Yd, Xd)
        assert np.allclose(backend.to_host(Yd), OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg, adjoint=True),
                           rtol=1e-14, atol=1e-14)
    n = (10, 4, 25)
    H = F.Helmholtz3D(*n, OO.helmholtz_diagonal(), -1.0, backend=backend)
    H.apply(Yd, Xd)
    assert np.allclose(backend.to_host(Yd), OO.helmholtz_apply(X, n, OO.helmholtz_diagonal()), rtol=1e-14,
                       atol=1e-14)
    H.adjoint().apply(Yd, Xd)
    assert np.allclose(backend.to_host(Yd), OO.helmholtz_apply(X, n, OO.helmholtz_diagonal(), adjoint=True),
                       rtol=1e-14, atol=1e-14)
    assert F.check_operator(H, seed=1)


@pytest.mark.parametrize("n,w", [((216, 216, 216), 2), ((37, 11, 5), 3), ((1, 1, 7), 2), ((5, 1, 1), 1),
                                 ((2, 3, 50), 2), ((600, 3, 4), 2), ((16, 70, 3), 2), ((65, 9, 4), 2),
                                 ((100, 45, 7), 1), ((17, 1, 30), 2), ((20, 2, 1), 2), ((15, 40, 3), 1)])
def test_helmholtz_rows_at_size(backend, n, w):
    """The plane-blocked stencil at the benchmark's 216^3 grid, on odd grids,
    on grids with an edge of 1 or 2, over several groups of planes (n3 = 30,
    50); on 2-D plane tiles from n1 = 16 (a ragged last tile in i (65) and in
    j (16 x 70, 100 x 45), one or two rows (n2 = 1, 2), one plane), and below
    n1 = 16 on 1-D tiles; in views with leading dimensions above N."""
    import funicular_hip as F
    rng = np.random.default_rng(20)
    N = n[0] * n[1] * n[2]
    X = randmatrix(rng, N, w, np.complex128)
    pad = 0 if N > 10 ** 6 else 13
    Xd = backend.todevice(np.asfortranarray(np.vstack([X, np.zeros((pad, w))])))
    Yd = backend.alloc_device(np.complex128, (N + pad, w))
    H = F.Helmholtz3D(*n, OO.helmholtz_diagonal(), -1.0, backend=backend)
    for adj in (False, True):
        Hop = H.adjoint() if adj else H
        Hop.apply(Yd.view((0, N), None), Xd.view((0, N), None))
        got = backend.to_host(Yd)[:N]
        ref = OO.helmholtz_apply(X, n, OO.helmholtz_diagonal(), adjoint=adj)
        assert np.allclose(got, ref, rtol=1e-14, atol=1e-14)
    X32 = X.astype(np.complex64)
    Xd32 = backend.todevice(np.asfortranarray(X32))
    Yd32 = backend.alloc_device(np.complex64, (N, w))
    with pytest.raises(F.ArgumentError):
        H.apply(Yd32, Xd32)
    F.Helmholtz3D(*n, OO.helmholtz_diagonal(), -1.0, backend=backend, dtype=np.complex64).apply(Yd32, Xd32)
    ref = OO.helmholtz_apply(X32.astype(np.complex128), n, OO.helmholtz_diagonal())
    assert np.allclose(backend.to_host(Yd32), ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("N", [1, 2, 257, 3 * 2 ** 20 + 5])
def test_tridiag_rows_at_size(backend, N):
    import funicular_hip as F
    rng = np.random.default_rng(21)
    w = 3
    X = randmatrix(rng, N, w, np.complex128)
    diag = 1 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    Xd = backend.todevice(X)
    Yd = backend.alloc_device(np.complex128, (N, w))
    for dg in (None, diag):
        G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, diag=dg, backend=backend)
        G.apply(Yd, Xd)
        assert np.allclose(backend.to_host(Yd), OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg), rtol=1e-14, atol=1e-14)
        G.adjoint().apply(Yd, Xd)
        assert np.allclose(backend.to_host(Yd), OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg, adjoint=True),
                           rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("m,k", [(257, 33), (1000, 64), (3001, 100), (2048, 200)])
@pytest.mark.parametrize("stored", [np.complex64, np.complex128])
def test_fp32_products_over_tile_widths(backend, m, k, stored):
    """The FP32 Gram and solve on their swizzled 64-wide LDS planes, at widths
    that leave a ragged last tile and column block: native ComplexF32 data and
    FP32 products of ComplexF64 data."""
    rng = np.random.default_rng(m + k)
    A = randmatrix(rng, m, k, stored)
    Ad = backend.todevice(A)
    C = backend.alloc_device(np.complex128 if stored == np.complex128 else np.complex64, (k, k))
    backend.gram_accumulate(C, Ad, Ad, 1.0, 0.0, compute=np.complex64)
    H = backend.to_host(C)
    A64 = A.astype(np.complex128)
    ref = A64.conj().T @ A64
    assert np.array_equal(H, H.conj().T)
    assert np.linalg.norm(H - ref) <= 1e-5 * np.linalg.norm(ref)
    # The solve by R of the FP64 Gram, in FP32 products.
    R = np.asfortranarray(np.linalg.cholesky(ref).conj().T)
    Yd = backend.todevice(A)
    backend.rdiv_upper(Yd, backend.todevice(R.astype(stored)), compute=np.complex64)
    X = backend.to_host(Yd).astype(np.complex128)
    Xr = OB.rdiv_upper(A64, R)
    assert np.linalg.norm(X - Xr) <= 1e-4 * np.linalg.norm(Xr)
    assert np.linalg.norm(X.conj().T @ X - np.eye(k)) <= 1e-3


def test_ghost_fill_matches_the_fixture(backend):
    """The device fill against tests/golden/ghost_stream.json (the oracle's
    stream, committed): rows from 0 and from 1,000,000 of columns 1, 2 and 300
    under two seeds, ComplexF64 and Float64, bit for bit."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ghost_stream.json")
    fx = json.load(open(path))
    for case in fx["cases"]:
        n = case["rows"]
        for col in fx["cols"]:
            d = backend.alloc_device(np.complex128, (n, 1))
            backend.ghost_fill(d, case["row0"], col, case["seed"])
            want = np.array([complex(a, b) for a, b in case["complex128"][str(col)]])
            assert np.array_equal(backend.to_host(d)[:, 0], want)
            r = backend.alloc_device(np.float64, (n, 1))
            backend.ghost_fill(r, case["row0"], col, case["seed"])
            assert np.array_equal(backend.to_host(r)[:, 0], np.array(case["float64"][str(col)]))
