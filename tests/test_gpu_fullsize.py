"""The 216^3 x 512 Helmholtz basis (N = 10,077,696, the benchmark's size),
resident in HBM, through size-independent properties: the QR invariants
Q^H Q = I and Q R = Y (Y regenerated panel by panel from its ghost seed), the
mixed first pass's orthogonality bar, and S - S^H = 2 i Im(d) Q^H Q for the
projected Helmholtz operator G = d I + (real symmetric stencil)."""
import gc

import numpy as np
import pytest

import funicular_hip as F

pytestmark = pytest.mark.gpu

N_EDGE, K, W = 216, 512, 32
SEED = 0xDEADBEEF
D = 6.0 - 0.09 * (1 + 0.05j)


@pytest.fixture(scope="module")
def basis(backend):
    N = N_EDGE ** 3
    plan = F.ResidencyPlan(backend=backend, device_budget=200 * 1024 ** 3, host_budget=8 * 1024 ** 3,
                           device_resident=True)
    G = F.Helmholtz3D(N_EDGE, N_EDGE, N_EDGE, D, -1.0, backend=backend)
    omega = F.GhostPanels(F.device_randn, np.complex128, N, K, plan=plan, seed=SEED, w=W)
    Y = F.PanelMatrix(N, K, plan=plan, w=W)
    yield plan, G, omega, Y
    F.free(Y)
    del plan, G, omega, Y
    gc.collect()


def residual(backend, Y, R, G):
    """||Y0 - Q R||_F / ||Y0||_F with Y0 = G Omega regenerated a panel at a time."""
    N, k, w = Y.N, Y.k, Y.w
    Rd = backend.todevice(np.asfortranarray(R))
    y0, qr, om = (backend.alloc_device(np.complex128, (N, w)) for _ in range(3))
    g = backend.alloc_device(np.complex128, (w, w))
    num = den = 0.0
    for j in range(0, k, w):
        backend.ghost_fill(om, 0, j + 1, SEED)
        G.apply(y0, om)
        backend.rightmul_gemm(qr, Y.device.view(None, (0, j + w)), Rd.view((0, j + w), (j, j + w)))
        backend.gram_accumulate(g, y0, y0, 1.0, 0.0)
        den += float(np.trace(backend.to_host(g)).real)
        backend.axpby(qr, y0, -1.0, 1.0)
        backend.gram_accumulate(g, qr, qr, 1.0, 0.0)
        num += float(np.trace(backend.to_host(g)).real)
    return float(np.sqrt(num / den))


def test_cholqr2_at_full_size(backend, basis):
    plan, G, omega, Y = basis
    print("full size: fp64 cholqr2", flush=True)
    F.panelmul(Y, G, omega)
    R = F.cholqr2(Y)
    assert np.allclose(R, np.triu(R)) and (np.diag(R).real > 0).all()
    orth = float(np.linalg.norm(F.gram(Y) - np.eye(K)))
    assert orth <= 1e-13, orth
    res = residual(backend, Y, R, G)
    assert res <= 1e-13, res
    # S - S^H = 2 i Im(d) Q^H Q for the orthonormal Q.
    S = F.project(Y, G)
    err = float(np.linalg.norm(S - S.conj().T - 2j * D.imag * np.eye(K)) / np.linalg.norm(S))
    assert err <= 1e-13, err


def test_mixed_cholqr2_meets_the_orthogonality_bar(backend, basis):
    plan, G, omega, Y = basis
    print("full size: mixed cholqr2", flush=True)
    for solve in ("fp32", "fp64"):
        F.panelmul(Y, G, omega)
        R = F.cholqr2_mixed(Y, solve=solve)
        orth = float(np.linalg.norm(F.gram(Y) - np.eye(K)))
        assert orth <= 1e-13, (solve, orth)
        res = residual(backend, Y, R, G)
        # The FP32 solve leaves the factorization at FP32 level; the FP64
        # solve keeps it at FP64 level.
        assert res <= (1e-5 if solve == "fp32" else 1e-13), (solve, res)


def test_sketch_chain_at_its_configured_size(backend):
    """The 1-D sketch chain at its configured size, N = 2^20, k = 64 in panels
    of 8: Omega = GhostPanels(seed 0xdeadbeef), Y = G Omega for the periodic
    tridiagonal (2 + 0.1i, -1), cholqr2(Y), S = project(Q, G), against the
    oracle's chain on the same Omega (the ghost stream itself is bit-pinned to
    the oracle at smaller sizes). Bars: ||Q^H Q - I||_F <= 1e-13,
    ||Y - Q R||_F / ||Y||_F <= 1e-13, S within 1e-12 normwise and in
    every entry, relative to that entry."""
    from oracle import operators as OO
    from oracle import panelblas as OB
    print("full size: the 2^20 x 64 sketch chain", flush=True)
    N, k, w = 2 ** 20, 64, 8
    d, o = 2 + 0.1j, -1.0
    plan = F.ResidencyPlan(backend=backend, device_budget=16 * 1024 ** 3, host_budget=8 * 1024 ** 3)
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=SEED, w=w)
    G = F.Tridiagonal1D(N, d, o, backend=backend)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, omega)
    Om = F.to_matrix(omega)
    Y0 = F.to_matrix(Y)
    R = F.cholqr2(Y)
    S = F.project(Y, G)
    Q = F.to_matrix(Y)
    apply = lambda X: OO.tridiag_apply(X, d, o)
    Yo = OB.panelmul(apply, Om, w)
    assert OB.isapprox(Y0, Yo, 1e-14)
    Qo, Ro = OB.cholqr2(Yo, w)
    So = OB.project(Qo, apply, w)
    assert OB.orthogonality_error(Q) <= 1e-13
    assert np.linalg.norm(Q @ R - Yo) / np.linalg.norm(Yo) <= 1e-13
    assert OB.isapprox(R, Ro, 1e-12)
    assert np.linalg.norm(S - So) <= 1e-12 * np.linalg.norm(So)
    # Every entry, the smallest (|S_ij| ~ 3e-5 of max 3.3) included, within
    # 1e-12 relative of the oracle's.
    rel = np.abs(S - So) / np.abs(So)
    print(f"S entrywise relative: max {rel.max():.2e}, |S_ij| from {np.abs(So).min():.2e} to "
          f"{np.abs(So).max():.2e}", flush=True)
    assert np.all(rel <= 1e-12)
    F.free(Y)
    del omega, plan
    gc.collect()
