"""cholqr2 with the mixed-precision first pass on Y = U diag(sigma) V^H with
log-spaced singular values (the mixed-precision workload at test size):
orthogonality at FP64 level, R against the FP64 cholqr2's, and Q R against
Y, with the solve in FP64 and in FP32."""
import numpy as np
import pytest
import scipy.linalg

import funicular_hip as F
from funicular_hip.panelblas import cholesky_upper_fp32, cholqr2_mixed
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu


def graded(rng, N, k, decades):
    U = scipy.linalg.qr(rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k)), mode="economic")[0]
    V = scipy.linalg.qr(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k)))[0]
    sigma = 10.0 ** (-decades * np.arange(k) / (k - 1))
    return np.asfortranarray((U * sigma) @ V.conj().T)


def make_plan(backend):
    return F.ResidencyPlan(backend=backend, device_budget=256 * 2 ** 20, host_budget=256 * 2 ** 20)


@pytest.mark.parametrize("decades", (2, 4))
@pytest.mark.parametrize("solve", ("fp64", "fp32"))
@pytest.mark.parametrize("w", (16, 7))
def test_cholqr2_mixed_on_graded_spectrum(backend, decades, solve, w):
    rng = np.random.default_rng(60 + decades)
    N, k = 20000, 64
    Y0 = graded(rng, N, k, decades)
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    R = cholqr2_mixed(Y, solve=solve)
    Q = F.to_matrix(Y)
    assert np.allclose(R, np.triu(R))
    assert np.all(np.diag(R).real > 0)
    assert OB.orthogonality_error(Q) <= 1e-13
    resid = np.linalg.norm(Q @ R - Y0) / np.linalg.norm(Y0)
    if solve == "fp64":
        assert resid <= 1e-13
    else:
        # The FP32 solve leaves an error of order u32 kappa(Y) in Q R.
        assert resid <= 1e-7 * 10.0 ** decades
    Qo, Ro = OB.cholqr2(Y0, w)
    # R agrees with the FP64 factor to the conditioning of the problem.
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= (1e-14 if solve == "fp64" else 1e-6) * 10.0 ** decades


def test_cholqr2_mixed_matches_fp64_on_well_conditioned(backend):
    rng = np.random.default_rng(61)
    N, k, w = 30000, 40, 8
    Y0 = np.asfortranarray(rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k)))
    plan = make_plan(backend)
    Y = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    R = cholqr2_mixed(Y, solve="fp64")
    Q = F.to_matrix(Y)
    Z = F.PanelMatrix.from_array(Y0, plan=plan, w=w)
    R64 = F.cholqr2(Z)
    assert OB.isapprox(R, R64, 1e-13)
    assert OB.isapprox(Q, F.to_matrix(Z), 1e-13)


def test_fp32_cholesky_shift_policy(backend):
    # Host-side rule, exercised here beside its kernels: "adaptive" shifts only
    # when the unshifted FP32 factorization fails; "fukaya" always shifts.
    rng = np.random.default_rng(62)
    A = rng.standard_normal((500, 20)) + 1j * rng.standard_normal((500, 20))
    G = A.conj().T @ A
    R, s = cholesky_upper_fp32(G, 500, 20)
    assert s == 0.0 and R.dtype == np.complex64
    R, s = cholesky_upper_fp32(G, 500, 20, shift="fukaya")
    lam = np.linalg.eigvalsh(G)[-1]
    assert s == pytest.approx(11.0 * (500 * 20 + 20 * 21) * 2.0 ** -24 * lam, rel=1e-12)
    # A singular Gram: the unshifted factorization fails and the shift takes over.
    B = A.copy()
    B[:, 3] = B[:, 2]
    R, s = cholesky_upper_fp32(B.conj().T @ B, 500, 20)
    assert s > 0.0
    with pytest.raises(F.ArgumentError):
        cholqr2_mixed(F.PanelMatrix.from_array(np.asfortranarray(A), plan=make_plan(backend), w=5), solve="fp16")


def test_mixed_workload_at_its_configured_size(backend):
    """The mixed-precision workload at its configured size: Y = U diag(sigma)
    V^H, 8,388,608 x 256 in panels of 16, resident in HBM, U the Q of cholqr2
    of a ghost basis (seed 0xdeadbeef), V a seeded 256 x 256 unitary, sigma
    log-spaced to 1e-2 and to 1e-4. The mixed first pass (FP32 Gram,
    Cholesky and solve, or the solve in FP64), with the adaptive shift and
    with the stated shift 11 (N k + k (k+1)) u32 ||Y||_2^2 always applied, is
    held to the FP64-only run: ||Q^H Q - I||_F <= 1e-13 for all, ||Y - Q R|| /
    ||Y|| <= 1e-13 with the FP64 solves and <= 1e-7 kappa with the FP32 one,
    and R within the conditioning of the FP64-only R. The always-shifted
    runs take three passes (six row sweeps), the others two."""
    from funicular_hip.executor import readpanels, sweep, trace_sweeps, updatepanels, writepanels
    print("full size: the 8388608 x 256 mixed workload", flush=True)
    N, k, w = 8388608, 256, 16
    plan = F.ResidencyPlan(backend=backend, device_budget=200 * 1024 ** 3, host_budget=1 << 30,
                           device_resident=True)
    U = F.PanelMatrix(N, k, plan=plan, w=w)
    src = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(src), writepanels(U))
    F.cholqr2(U)
    rng = np.random.default_rng(0xDEADBEEF + 1)
    V = np.linalg.qr(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k)))[0]
    Y0, Y, Z = (F.PanelMatrix(N, k, plan=plan, w=w) for _ in range(3))

    def resid(R):
        F.rightmul(Z, Y, R)
        sweep(lambda _, a, b: backend.axpby(b, a, -1.0, 1.0), readpanels(Y0), updatepanels(Z))
        return float(np.sqrt(np.real(np.trace(F.gram(Z))) / np.real(np.trace(F.gram(Y0)))))

    try:
        for decades in (2, 4):
            kappa = 10.0 ** decades
            sigma = 10.0 ** (-decades * np.arange(k) / (k - 1))
            F.rightmul(Y0, U, sigma[:, None] * V.conj().T)
            results = {}
            for name, fn in (("fp64", F.cholqr2), ("mixed", cholqr2_mixed),
                             ("mixed_fp64_solve", lambda X: cholqr2_mixed(X, solve="fp64")),
                             ("fukaya", lambda X: cholqr2_mixed(X, shift="fukaya")),
                             ("fukaya_fp64_solve", lambda X: cholqr2_mixed(X, shift="fukaya", solve="fp64"))):
                sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(Y0), writepanels(Y))
                with trace_sweeps() as sweeps:
                    R = fn(Y)
                orth = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
                results[name] = (R, orth, resid(R), len(sweeps))
            R64 = results["fp64"][0]
            for name, (R, orth, res, nsweeps) in results.items():
                fp32_solve = name in ("mixed", "fukaya")
                print(f"sigma to 1e-{decades} {name}: orthogonality {orth:.2e}, residual {res:.2e}, "
                      f"{nsweeps} sweeps", flush=True)
                assert nsweeps == (6 if name.startswith("fukaya") else 4), (name, nsweeps)
                assert orth <= 1e-13, (name, orth)
                assert res <= (1e-7 * kappa if fp32_solve else 1e-13), (name, res)
                tol = (1e-6 if fp32_solve else 1e-14) * kappa
                assert np.linalg.norm(R - R64) <= tol * np.linalg.norm(R64), name
    finally:
        for pm in (U, Y0, Y, Z):
            F.free(pm)
        del src, plan
        import gc
        gc.collect()
