"""The adjoints of cholqr2 and project on the HIP backend: the kernels they
add (Y R^-H, the row-wise real dot product, Y <- aX + bY) against dense
references, the pullbacks against the oracle's, and both against central
finite differences of the device forward pass."""
import numpy as np
import pytest
import scipy.linalg

import funicular_hip as F
from oracle import adjoint as OA
from oracle import operators as OO
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

C16 = np.complex128
EL = [np.complex128, np.complex64, np.float64, np.float32]


def randmatrix(rng, N, k, dtype=C16):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(dtype, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(dtype))


def make_plan(backend, **kw):
    kw.setdefault("device_budget", 64 * 2 ** 20)
    kw.setdefault("host_budget", 64 * 2 ** 20)
    return F.ResidencyPlan(backend=backend, **kw)


def tol(T):
    return 1e-4 if np.dtype(T) in (np.dtype(np.complex64), np.dtype(np.float32)) else 1e-12


@pytest.mark.parametrize("T", EL)
@pytest.mark.parametrize("m,k", [(7, 5), (1000, 64), (333, 45), (2000, 96), (100, 1), (4100, 130)])
def test_rdiv_upper_adjoint(backend, T, m, k):
    rng = np.random.default_rng(m + 7 * k)
    Y = randmatrix(rng, m, k, T)
    # A well-conditioned factor: the Cholesky factor of a tall random matrix's Gram.
    B = randmatrix(rng, 2 * k + 10, k, np.complex128 if np.issubdtype(T, np.complexfloating) else np.float64)
    R = np.asfortranarray(np.linalg.cholesky(B.conj().T @ B).conj().T.astype(T))
    Yd = backend.todevice(Y)
    backend.rdiv_upper_adjoint(Yd, backend.todevice(R))
    X = backend.to_host(Yd)
    ref = scipy.linalg.solve_triangular(R.astype(np.complex128), Y.astype(np.complex128).conj().T,
                                        lower=False).conj().T
    assert np.linalg.norm(X - ref) <= tol(T) * np.linalg.norm(ref)
    # X R^H recovers Y.
    assert np.linalg.norm(X @ R.conj().T - Y) <= tol(T) * np.linalg.norm(Y) * 10


def test_rdiv_upper_adjoint_view_leaves_other_rows(backend):
    rng = np.random.default_rng(3)
    Y = randmatrix(rng, 300, 40)
    R = np.asfortranarray(np.triu(randmatrix(rng, 40, 40)) + 4 * np.eye(40))
    S = backend.todevice(Y)
    backend.rdiv_upper_adjoint(S.view((0, 170), None), backend.todevice(R))
    H = backend.to_host(S)
    ref = scipy.linalg.solve_triangular(R, Y[:170].conj().T, lower=False).conj().T
    assert OB.isapprox(H[:170], ref, 1e-12)
    assert np.array_equal(H[170:], Y[170:])


@pytest.mark.parametrize("T", EL)
def test_rowdot_and_axpby(backend, T):
    rng = np.random.default_rng(5)
    A, B, C = (randmatrix(rng, 1001, 9, T) for _ in range(3))
    out = backend.alloc_device(np.float64, (1001, 1))
    backend.rowdot(out, backend.todevice(A), backend.todevice(B))
    ref = np.real(np.sum(A.conj().astype(C16) * B.astype(C16), axis=1))
    assert np.allclose(backend.to_host(out).ravel(), ref, rtol=tol(T), atol=tol(T))
    backend.rowdot(out, backend.todevice(C), backend.todevice(B), accumulate=True)
    ref = ref + np.real(np.sum(C.conj().astype(C16) * B.astype(C16), axis=1))
    assert np.allclose(backend.to_host(out).ravel(), ref, rtol=tol(T), atol=tol(T))
    Yd = backend.todevice(A)
    a, b = (0.5 - 0.25j, 2.0 + 1j) if np.issubdtype(T, np.complexfloating) else (0.5, 2.0)
    backend.axpby(Yd, backend.todevice(B), a, b)
    assert np.allclose(backend.to_host(Yd), a * B + b * A, rtol=tol(T), atol=tol(T))


@pytest.mark.parametrize("w,resident", [(16, True), (7, False), (16, False)])
def test_cholqr2_pullback_against_oracle(backend, w, resident):
    rng = np.random.default_rng(40 + w)
    N, k = 4096, 16
    Y0 = randmatrix(rng, N, k)
    Qb, Rb = randmatrix(rng, N, k), np.triu(randmatrix(rng, k, k))
    plan = make_plan(backend, device_resident=resident) if resident else make_plan(backend)
    Y = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    R, pullback = F.rrule_cholqr2(Y)
    Q = F.to_matrix(Y)
    Qbar = F.PanelMatrix.from_array(Qb, plan=plan, w=w)
    Ybar = F.to_matrix(pullback(Qbar, Rb))
    ref = OA.qr_pullback(Q, R, Qb, Rb)
    assert OB.isapprox(Ybar, ref, 1e-11)
    # Each cotangent alone.
    assert OB.isapprox(F.to_matrix(pullback(None, Rb)), OA.qr_pullback(Q, R, 0 * Qb, Rb), 1e-11)
    assert OB.isapprox(F.to_matrix(pullback(Qbar, None)), OA.qr_pullback(Q, R, Qb, 0 * Rb), 1e-11)


def test_cholqr2_pullback_finite_differences(backend):
    # The device forward pass differentiated numerically, direction by direction.
    rng = np.random.default_rng(41)
    N, k, w = 4096, 16, 8
    Y0 = randmatrix(rng, N, k)
    Qb, Rb = randmatrix(rng, N, k), np.triu(randmatrix(rng, k, k))
    plan = make_plan(backend)

    def loss(Yh):
        Y = F.PanelMatrix.from_array(np.asfortranarray(Yh), plan=plan, w=w)
        R = F.cholqr2(Y)
        return np.real(np.vdot(Qb, F.to_matrix(Y)) + np.vdot(Rb, R))

    Y = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    _, pullback = F.rrule_cholqr2(Y)
    Ybar = F.to_matrix(pullback(F.PanelMatrix.from_array(Qb, plan=plan, w=w), Rb))
    h = 1e-6
    for _ in range(3):
        E = randmatrix(rng, N, k)
        fd = (loss(Y0 + h * E) - loss(Y0 - h * E)) / (2 * h)
        assert abs(fd - np.real(np.vdot(Ybar, E))) <= 1e-6 * max(1.0, abs(fd))


def test_cholqr2_pullback_real(backend):
    rng = np.random.default_rng(42)
    N, k, w = 3000, 12, 5
    Y0 = randmatrix(rng, N, k, np.float64)
    Qb, Rb = randmatrix(rng, N, k, np.float64), np.triu(randmatrix(rng, k, k, np.float64))
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    R, pullback = F.rrule_cholqr2(Y)
    Q = F.to_matrix(Y)
    Ybar = F.to_matrix(pullback(F.PanelMatrix.from_array(Qb, plan=plan, w=w), Rb))
    assert Ybar.dtype == np.float64
    assert OB.isapprox(Ybar, np.real(OA.qr_pullback(Q, R, Qb, Rb)), 1e-11)


@pytest.mark.parametrize("w,resident", [(16, False), (5, False), (16, True)])
def test_project_pullback_against_oracle(backend, w, resident):
    rng = np.random.default_rng(50 + w)
    N, k = 4096, 16
    d0, o0 = 2 + 0.1j, -1.0
    diag = 1 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    G = F.Tridiagonal1D(N, d0, o0, diag, backend=backend)
    Qh = randmatrix(rng, N, k)
    Sb = randmatrix(rng, k, k)
    plan = make_plan(backend, device_resident=True) if resident else make_plan(backend)
    Q = F.PanelMatrix.from_array(Qh, plan=plan, w=w)
    S, pullback = F.rrule_project(Q, G)
    Qbar, dbar = pullback(Sb)
    if resident:
        # A second call reuses the pooled scratch and gives the same bits.
        Qbar2, dbar2 = pullback(Sb)
        assert np.array_equal(F.to_matrix(Qbar2), F.to_matrix(Qbar)) and np.array_equal(dbar2, dbar)
    Qr, dr = OA.project_pullback(Qh, lambda X: OO.tridiag_apply(X, d0, o0, diag),
                                 lambda X: OO.tridiag_apply(X, d0, o0, diag, adjoint=True), Sb)
    assert OB.isapprox(F.to_matrix(Qbar), Qr, 1e-12)
    assert np.allclose(dbar, dr, rtol=1e-12, atol=1e-12 * np.abs(dr).max())
    assert np.array_equal(F.to_matrix(Q), Qh)

    # Central differences of the device forward pass, in Q and in the diagonal.
    def loss(Qh2, diag2):
        G2 = F.Tridiagonal1D(N, d0, o0, diag2, backend=backend)
        return np.real(np.vdot(Sb, F.project(F.PanelMatrix.from_array(np.asfortranarray(Qh2), plan=plan, w=w), G2)))
    h = 1e-6
    E = randmatrix(rng, N, k)
    fd = (loss(Qh + h * E, diag) - loss(Qh - h * E, diag)) / (2 * h)
    assert abs(fd - np.real(np.vdot(F.to_matrix(Qbar), E))) <= 1e-6 * max(1.0, abs(fd))
    e = rng.standard_normal(N)
    fd = (loss(Qh, diag + h * e) - loss(Qh, diag - h * e)) / (2 * h)
    assert abs(fd - dbar @ e) <= 1e-6 * max(1.0, abs(fd))


def test_adjoint_workload_at_its_configured_size(backend):
    """The adjoint workload at its configured size: Y a 4,194,304 x 128 ghost
    basis in panels of 32, resident in HBM; Q R = Y and S = Q^H G Q with
    G = diag(d) + the periodic tridiagonal (2 + 0.1i, -1), d_i = 1 + 0.5
    sin(2 pi i / N); loss Re<dQ, Q> + Re<dS, S>. Both pullbacks are checked
    against central differences of the device forward pass, along a ghost
    direction E of Y and a random direction e of d."""
    from funicular_hip.executor import readpanels, sweep, updatepanels, writepanels
    print("full size: the 4194304 x 128 adjoint workload", flush=True)
    N, k, w = 1 << 22, 128, 32
    plan = F.ResidencyPlan(backend=backend, device_budget=100 * 1024 ** 3, host_budget=1 << 30,
                           device_resident=True)
    d = 1.0 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, d, backend=backend)

    def ghost_copy(seed):
        src = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=seed, w=w)
        dst = F.PanelMatrix(N, k, plan=plan, w=w)
        sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(src), writepanels(dst))
        return dst

    rng = np.random.default_rng(4)
    A = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    dS = (A + A.conj().T) / 2
    dQ = ghost_copy(0xDEADBEEF + 4)

    def loss(Y, S):
        return float(np.real(F.gram(dQ, Y).trace()) + np.real(np.vdot(dS, S)))

    Y = ghost_copy(0xDEADBEEF + 3)
    R, pb_qr = F.rrule_cholqr2(Y)
    S, pb_pr = F.rrule_project(Y, G)
    Qbar, dbar = pb_pr(dS)
    sweep(lambda _, a, b: backend.axpby(b, a, 1.0, 1.0), readpanels(dQ), updatepanels(Qbar))
    Ybar = pb_qr(Qbar, None)
    E = ghost_copy(0xDEADBEEF + 5)
    h = 1e-6

    def shifted(sign):
        Yh = ghost_copy(0xDEADBEEF + 3)
        sweep(lambda _, e, y: backend.axpby(y, e, sign * h, 1.0), readpanels(E), updatepanels(Yh))
        F.cholqr2(Yh)
        out = loss(Yh, F.project(Yh, G))
        F.free(Yh)
        return out

    fd = (shifted(1.0) - shifted(-1.0)) / (2 * h)
    directional = float(np.real(F.gram(Ybar, E).trace()))
    rel = abs(fd - directional) / abs(fd)
    e = np.random.default_rng(5).standard_normal(N)

    def shifted_d(sign):
        Gh = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, d + sign * h * e, backend=backend)
        return float(np.real(np.vdot(dS, F.project(Y, Gh))))

    fd_d = (shifted_d(1.0) - shifted_d(-1.0)) / (2 * h)
    rel_d = abs(fd_d - float(dbar @ e)) / abs(fd_d)
    print(f"adjoint at full size: <Ybar, E> rel {rel:.2e}; dS/dd along e rel {rel_d:.2e}", flush=True)
    assert rel <= 1e-6, rel
    assert rel_d <= 1e-5, rel_d
    for pm in (Y, dQ, Qbar, Ybar, E):
        F.free(pm)
    del R, S, pb_qr, pb_pr, plan
    import gc
    gc.collect()
