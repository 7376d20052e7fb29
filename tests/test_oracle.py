"""Pins the oracle itself: against brute force, closed forms, textbook
routines and the reference's known answers, so that a dropped term, a wrong
sign or a transposed operand in the oracle fails here."""
import math

import mpmath
import numpy as np
import pytest
import scipy.linalg
from scipy import stats

from oracle import ghost as OG
from oracle import operators as OO
from oracle import panelblas as OB
from oracle import plan as OP


def randmatrix(rng, N, k, dtype=np.complex128):
    """test/setup.jl:148: uniform in [-1/2, 1/2] (both parts for complex)."""
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(dtype, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return A.astype(dtype)


def test_gram_matches_mpmath_brute_force():
    rng = np.random.default_rng(1)
    A = randmatrix(rng, 9, 4)
    B = randmatrix(rng, 9, 3)
    G = OB.gram(A, B, w=2)
    for i in range(4):
        for j in range(3):
            exact = mpmath.fsum(mpmath.conj(mpmath.mpc(A[r, i])) * mpmath.mpc(B[r, j]) for r in range(9))
            assert abs(complex(exact) - G[i, j]) < 1e-14
    # A^H B, not A^T B nor B^H A: a transposed or unconjugated product fails.
    assert not np.allclose(G, (A.T @ B))
    assert not np.allclose(G, (B.conj().T @ A).T)


def test_gram_sums_every_row_block():
    rng = np.random.default_rng(2)
    A = randmatrix(rng, 40, 23)
    for w in (23, 12, 7, 1):
        G = OB.gram(A, w=w)
        assert OB.isapprox(G, A.conj().T @ A, OB.blastol(np.complex128, 40, 23))
    # A ragged last block (N not a multiple of h) is still summed.
    assert OP.row_blocks(40, OP.row_block_height(40, 23, 7))[-1][1] == 40


def test_cholesky_shift_known_answers():
    # test/test_panelblas.jl:327-345
    rng = np.random.default_rng(3)
    N, k = 40, 23
    H = randmatrix(rng, N, k)
    G = H.conj().T @ H
    R, shifted = OB.cholesky_upper(G, N, k)
    assert not shifted
    assert np.allclose(R.conj().T @ R, G)
    assert np.allclose(R, np.triu(R)) and np.all(np.diag(R).real > 0)
    S = G.copy()
    S[:, -1] = 0
    S[-1, :] = 0
    R, shifted = OB.cholesky_upper(S, N, k)
    assert shifted
    assert np.allclose(R, np.triu(R))
    s = 11 * N * k * np.finfo(float).eps * np.linalg.norm(S)
    assert np.allclose(R.conj().T @ R, S + s * np.eye(k))
    with pytest.raises(ValueError):
        OB.cholesky_upper(-G, N, k)


def test_rdiv_is_the_right_upper_solve():
    rng = np.random.default_rng(4)
    Y = randmatrix(rng, 7, 5)
    R = np.triu(randmatrix(rng, 5, 5)) + 4 * np.eye(5)
    X = OB.rdiv_upper(Y, R)
    assert np.allclose(X @ R, Y)
    assert not np.allclose(np.linalg.solve(R, Y.T).T, X)  # a left solve would differ


def test_cholqr2_invariants():
    rng = np.random.default_rng(5)
    A = randmatrix(rng, 40, 23)
    for w in (23, 12, 7, 1):
        Q, R = OB.cholqr2(A, w)
        assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(np.complex128, 40, 23)
        assert OB.isapprox(Q @ R, A, OB.blastol(np.complex128, 40, 23))
        assert np.allclose(R, np.triu(R))
        # Same R as a Householder QR up to the signs of the diagonal.
        Rh = scipy.linalg.qr(A, mode="r")[0][:23]
        ph = np.diag(Rh) / np.abs(np.diag(Rh))
        assert np.allclose(np.diag(ph.conj()) @ Rh, R, atol=1e-12)


def test_cholqr2_shifted_third_pass():
    # test/test_panelblas.jl:284-325: past kappa ~ eps^-1/2 the shift fires.
    rng = np.random.default_rng(6)
    N, k = 40, 23
    U = scipy.linalg.qr(randmatrix(rng, N, k), mode="economic")[0]
    V = scipy.linalg.qr(randmatrix(rng, k, k))[0]
    A = U @ np.diag(10.0 ** np.linspace(0, -10, k)) @ V.conj().T
    Q, R = OB.cholqr2(A, 7)
    assert OB.orthogonality_error(Q) <= 1e-8
    assert np.linalg.norm(Q @ R - A) <= 1e-10 * np.linalg.norm(A)


def test_project_against_explicit_product():
    rng = np.random.default_rng(7)
    N, k = 40, 23
    G = randmatrix(rng, N, N)
    Q = randmatrix(rng, N, k)
    for w in (23, 12, 7, 1):
        S = OB.project(Q, lambda X: G @ X, w)
        assert OB.isapprox(S, Q.conj().T @ G @ Q, OB.blastol(np.complex128, N, k))
        assert not np.isnan(S).any()  # every block assigned exactly once


def test_panelwise_cholqr_against_the_row_sweeps():
    # gram_panelwise is A^H A; rdiv_panelwise is Y R^-1 (a dropped update
    # term or a transposed block fails the residual); cholqr2_panelwise meets
    # the QR invariants and agrees with the row-sweep cholqr2.
    rng = np.random.default_rng(21)
    N, k = 200, 23
    A = randmatrix(rng, N, k)
    for w in (23, 12, 7, 1):
        G = OB.gram_panelwise(A, w)
        assert np.array_equal(np.tril(G, -1), np.triu(G, 1).conj().T)
        assert OB.isapprox(G, A.conj().T @ A, 1e-14)
        R, _ = OB.cholesky_upper(A.conj().T @ A, N, k)
        X = OB.rdiv_panelwise(A, R, w)
        assert np.linalg.norm(X @ R - A) <= 1e-13 * np.linalg.norm(A)
        Q, Rq = OB.cholqr2_panelwise(A, w)
        assert OB.orthogonality_error(Q) <= OB.orthogonality_bound(np.complex128, N, k)
        assert np.linalg.norm(Q @ Rq - A) <= 1e-13 * np.linalg.norm(A)
        assert np.allclose(Rq, np.triu(Rq)) and (np.diag(Rq).real > 0).all()
        Q2, R2 = OB.cholqr2(A, w)
        assert OB.isapprox(Rq, R2, 1e-12) and OB.isapprox(Q, Q2, 1e-12)
    # The shifted third pass: a condition number past eps^-1/2.
    U = scipy.linalg.qr(randmatrix(rng, N, k), mode="economic")[0]
    V = scipy.linalg.qr(randmatrix(rng, k, k))[0]
    B = U @ np.diag(10.0 ** np.linspace(0, -10, k)) @ V.conj().T
    Q, Rq = OB.cholqr2_panelwise(B, 7)
    assert OB.orthogonality_error(Q) <= 1e-8
    assert np.linalg.norm(Q @ Rq - B) <= 1e-10 * np.linalg.norm(B)


def test_affine_known_answers():
    # (1 + 2i)(3 - i) + (0.5 - 5i) = 5.5 exactly; 2 y + 3 on reals.
    assert OB.affine(np.array([1 + 2j]), 3 - 1j, 0.5 - 5j)[0] == 5.5 + 0j
    assert OB.affine(np.array([1.5, -2.0]), 2.0, 3.0).tolist() == [6.0, -1.0]
    # The product is rounded before the constant is added: (1 + 2^-27)^2 - 1
    # is 2^-26 that way, and 2^-26 + 2^-54 had the two been fused.
    t = 1.0 + 2.0 ** -27
    assert OB.affine(np.array([t]), t, -1.0)[0] == 2.0 ** -26
    assert OB.affine(np.array([t + 0j]), t, -1.0)[0] == complex(2.0 ** -26, 0.0)
    # alpha = 1, beta = 0 leaves the values; beta = 0 is scale.
    rng = np.random.default_rng(12)
    Y = rng.standard_normal(32) + 1j * rng.standard_normal(32)
    assert np.array_equal(OB.affine(Y, 1.0, 0.0), Y)
    assert np.array_equal(OB.affine(Y, 0.3 - 1.7j, 0.0), OB.scale(Y, 0.3 - 1.7j))


def test_scale_known_answers():
    # (1 + 2i)(3 - i) = 5 + 5i exactly; real scalars scale each part.
    assert OB.scale(np.array([1 + 2j]), 3 - 1j)[0] == 5 + 5j
    assert OB.scale(np.array([1.5 - 2j]), 2.0)[0] == 3 - 4j
    assert OB.scale(np.array([1.5, -2.0]), 2.0).tolist() == [3.0, -4.0]
    # Julia's Complex * Real never forms inf * 0, Complex * Complex does: a
    # dropped distinction between the two shows up here.
    y = np.array([complex(np.inf, 0.0)])
    assert OB.scale(y, 2.0)[0] == complex(np.inf, 0.0)
    z = OB.scale(y, 2.0 + 0j)[0]
    assert z.real == np.inf and np.isnan(z.imag)
    # Each part is the rounded sum of two rounded products (no fused step).
    rng = np.random.default_rng(11)
    Y = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    a = 0.3 - 1.7j
    got = OB.scale(Y, a)
    for yv, gv in zip(Y, got):
        re = float(np.float64(yv.real * a.real) - np.float64(yv.imag * a.imag))
        im = float(np.float64(yv.real * a.imag) + np.float64(yv.imag * a.real))
        assert gv.real == re and gv.imag == im


def test_dot_and_norm_match_exact_sums():
    rng = np.random.default_rng(12)
    N, k = 37, 11
    A = rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k))
    B = rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k))
    # Correctly rounded sums of the exact products (conj on the first operand).
    mp = mpmath.mp
    mp.prec = 200
    exact = mp.mpc(0)
    for a_, b_ in zip(A.ravel(), B.ravel()):
        exact += mp.conj(mp.mpc(a_.real, a_.imag)) * mp.mpc(b_.real, b_.imag)
    ref = complex(exact)
    sq = sum(mp.mpf(a_.real) ** 2 + mp.mpf(a_.imag) ** 2 for a_ in A.ravel())
    for w in (1, 4, 11):
        d = OB.dot(A, B, w)
        assert abs(d - ref) <= 1e-13 * abs(ref)
        assert abs(OB.norm(A, w) - float(mp.sqrt(sq))) <= 1e-14 * float(mp.sqrt(sq))
    # conj goes on the first operand: dot(A, iA) = i |A|^2.
    assert abs(OB.dot(A, 1j * A, 4) - 1j * float(sq)) <= 1e-13 * float(sq)
    assert OB.norm(np.array([[3.0, 4.0]]), 1) == 5.0
    # Real eltypes: a plain sum of products.
    Ar = A.real.copy()
    assert abs(OB.dot(Ar, Ar, 3) - float(np.sum(Ar * Ar))) <= 1e-13 * float(np.sum(Ar * Ar))


def test_axpy_is_alpha_x_plus_y():
    X = np.array([1 + 1j, 2 - 1j])
    Y = np.array([0.5j, -1.0])
    assert OB.axpy(2 - 1j, X, Y).tolist() == [(2 - 1j) * (1 + 1j) + 0.5j, (2 - 1j) * (2 - 1j) - 1.0]


def test_tridiagonal_against_dense():
    rng = np.random.default_rng(8)
    N = 17
    X = randmatrix(rng, N, 3)
    diag = 1 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    for dg in (None, diag):
        D = OO.tridiag_dense(N, 2 + 0.1j, -1.0, dg)
        assert np.allclose(OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg), D @ X)
        assert np.allclose(OO.tridiag_apply(X, 2 + 0.1j, -1.0, dg, adjoint=True), D.conj().T @ X)


def test_helmholtz_against_dense():
    rng = np.random.default_rng(9)
    n = (4, 3, 5)
    d = OO.helmholtz_diagonal()
    assert abs(d - (5.91 - 0.0045j)) < 1e-15
    X = randmatrix(rng, 60, 2)
    D = OO.helmholtz_dense(n, d)
    assert np.allclose(OO.helmholtz_apply(X, n, d), D @ X)
    assert np.allclose(OO.helmholtz_apply(X, n, d, adjoint=True), D.conj().T @ X)


def test_splitmix_published_outputs():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: its first outputs.
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC,
                0x1B39896A51A8749B]
    x = 0
    for e in expected:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        assert OG.mix(x) == e


def test_fixed_exp_and_log_track_libm():
    for v in np.linspace(-40, 0, 4001):
        assert abs(OG.exp(v) - math.exp(v)) <= 3e-16 * math.exp(v)
    for v in np.exp(np.linspace(-36.7, 0, 4001)):
        assert abs(OG.log(v) - math.log(v)) <= 4.5e-16 * max(1.0, abs(math.log(v)))


def test_ziggurat_layers_have_equal_area():
    x = OG.zig_edges()
    f = lambda t: math.exp(-0.5 * t * t)
    v = 4.92867323399e-3
    r = 3.6541528853610088
    tail = math.sqrt(math.pi / 2) * math.erfc(r / math.sqrt(2))
    assert abs(r * f(r) + tail - v) < 1e-11
    assert abs(x[0] * f(r) - v) < 1e-15
    for i in range(1, 255):
        assert abs(x[i] * (f(x[i + 1]) - f(x[i])) - v) < 1e-12
    assert 0 < x[255] < x[254]


def test_ghost_stream_is_standard_normal_and_keyed_by_column():
    A = OG.matrix(np.complex128, 100000, 4, 0xDEADBEEF)
    z = np.concatenate([A.real.ravel(), A.imag.ravel()]) * np.sqrt(2)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-4
    # Tail draws exist and are beyond r.
    assert np.abs(z).max() > 3.7
    # Width independence: any column range regenerates the same values.
    assert np.array_equal(OG.ghost(np.complex128, 0, 100000, 3, 2, 0xDEADBEEF), A[:, 2:4])
    assert np.array_equal(OG.ghost(np.complex128, 500, 7, 2, 1, 0xDEADBEEF), A[500:507, 1:2])
    assert not np.array_equal(OG.matrix(np.complex128, 100, 4, 0xDEADBEEF + 1), A[:100])
    R = OG.matrix(np.float64, 200000, 1, 7)
    assert abs(R.var() - 1) < 0.01


def _ghost_fixture():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ghost_stream.json")
    return json.load(open(path))


def test_ghost_stream_matches_its_fixture():
    # tests/golden/ghost_stream.json, written by tests/golden/make_ghost_fixture.py,
    # holds the stream's values; a change to any bit of it shows up here.
    fx = _ghost_fixture()
    for case in fx["cases"]:
        for col in fx["cols"]:
            z = OG.ghost(np.complex128, case["row0"], case["rows"], col, 1, case["seed"])[:, 0]
            want = np.array([complex(a, b) for a, b in case["complex128"][str(col)]])
            assert np.array_equal(z, want)
            x = OG.ghost(np.float64, case["row0"], case["rows"], col, 1, case["seed"])[:, 0]
            assert np.array_equal(x, np.array(case["float64"][str(col)]))
