"""End-to-end uses of the panel algebra on the HIP backend, each stage held
to dense arithmetic: the randomized factorization chain through a
rectangular operator (test/test_chain.jl:8-59) and the mini randomized SVD,
a subspace iteration ending in project (test/test_rsvd.jl:1-62). Tolerances
are the reference's: blastol = 100 eps sqrt(N k), orthogonality bound
100 N k eps (test/setup.jl:170-172)."""
import numpy as np
import pytest

import funicular_hip as F

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.complex128): 2.0 ** -52, np.dtype(np.complex64): 2.0 ** -23, np.dtype(np.float64): 2.0 ** -52}


def eps(T):
    return EPS[np.dtype(T)]


def blastol(T, N, k):
    return 100 * eps(T) * np.sqrt(N * k)


def orthogonality_bound(T, N, k):
    return 100 * N * k * eps(T)


def randmatrix(rng, T, m, n):
    A = rng.standard_normal((m, n))
    if np.issubdtype(np.dtype(T), np.complexfloating):
        A = A + 1j * rng.standard_normal((m, n))
    return np.asfortranarray(A.astype(T))


def close(a, b, rtol):
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def plan_for(backend):
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20)


@pytest.mark.parametrize("T", [np.complex128, np.complex64, np.float64])
def test_a_randomized_factorization_chain_matches_the_dense_reference(backend, T):
    m, n, s, r, w = 120, 96, 24, 9, 5
    rng = np.random.default_rng(0xDEADBEEF)
    plan = plan_for(backend)
    reference = randmatrix(rng, T, m, n)
    G = F.DenseOperator(reference, backend)
    tol = blastol(T, max(m, n), s)

    omega = F.GhostPanels(T, n, s, plan=plan, seed=0xDEADBEEF, w=w)
    probe = F.to_matrix(omega)
    Y = F.PanelMatrix(m, s, plan=plan, w=w, dtype=T)
    F.panelmul(Y, G, omega)
    sketch = reference.astype(np.complex128) @ probe
    assert close(F.to_matrix(Y), sketch, tol)

    R = F.cholqr2(Y)
    Q = F.to_matrix(Y).astype(np.complex128)
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(s), 2) <= orthogonality_bound(T, m, s)
    assert close(Q @ R, sketch, tol)

    Z = F.PanelMatrix(n, s, plan=plan, w=w, dtype=T)
    F.panelmul(Z, G.adjoint(), Y)
    image = reference.astype(np.complex128).conj().T @ Q
    assert close(F.to_matrix(Z), image, tol)

    S = F.cholqr2(Z)
    P = F.to_matrix(Z).astype(np.complex128)
    assert np.linalg.norm(P.conj().T @ P - np.eye(s), 2) <= orthogonality_bound(T, n, s)
    assert close(P @ S, image, tol)

    C = randmatrix(rng, T, s, r)
    V = F.PanelMatrix(n, r, plan=plan, w=w, dtype=T)
    F.rightmul(V, Z, C)
    assert close(F.to_matrix(V), P @ C, tol)

    collected = F.to_matrix(V)
    roundtrip = F.PanelMatrix.from_array(collected, plan=plan, w=w)
    assert np.array_equal(F.to_matrix(roundtrip), collected)
    wide = F.PanelMatrix(n, 2 * r, plan=plan, w=w, dtype=T)
    F.copycols(wide, range(0, r), collected)
    F.copycols(wide, range(r, 2 * r), roundtrip, range(0, r))
    assert np.array_equal(F.to_matrix(wide), np.hstack([collected, collected]))
    assert np.array_equal(F.to_matrix(omega), probe)


def lowrank_plus_noise(rng, T, N, rank, floor):
    B = randmatrix(rng, T, N, rank).astype(np.complex128)
    A = B @ B.conj().T
    A /= np.linalg.norm(A, 2)
    E = randmatrix(rng, T, N, N).astype(np.complex128)
    A = A + floor * (E + E.conj().T)
    A = (A + A.conj().T) / 2
    return np.asfortranarray(A.astype(T))


def minirsvd(backend, G, T, N, k, plan, w, rng, iterations=2, nbuffers=None):
    Q = F.PanelMatrix.from_array(randmatrix(rng, T, N, k), plan=plan, w=w)
    scratch = F.PanelMatrix(N, k, plan=plan, w=w, dtype=T)
    F.cholqr2(Q, nbuffers=nbuffers)
    for _ in range(iterations):
        F.panelmul(scratch, G, Q, nbuffers=nbuffers)
        F.cholqr2(scratch, nbuffers=nbuffers)
        Q, scratch = scratch, Q
    return Q, F.project(Q, G, nbuffers=nbuffers)


@pytest.mark.parametrize("w", (24, 12, 17, 1))
def test_mini_rsvd_matches_the_dense_reference(backend, w):
    T = np.complex128
    N, k, rank = 120, 24, 6
    rng = np.random.default_rng(0xDEADBEEF + w)
    A = lowrank_plus_noise(rng, T, N, rank, np.sqrt(eps(T)))
    reference = np.linalg.svd(A, compute_uv=False)
    Q, S = minirsvd(backend, F.DenseOperator(A, backend), T, N, k, plan_for(backend), w, rng)
    Qh = F.to_matrix(Q)
    assert np.linalg.norm(Qh.conj().T @ Qh - np.eye(k), 2) <= orthogonality_bound(T, N, k)
    got = np.linalg.svd(S, compute_uv=False)
    assert np.allclose(got[:rank], reference[:rank], rtol=1e-6, atol=0)


def test_mini_rsvd_is_unchanged_by_the_buffer_count(backend):
    T = np.complex128
    N, k, w, rank = 120, 24, 7, 6
    A = lowrank_plus_noise(np.random.default_rng(0xDEADBEEF), T, N, rank, np.sqrt(eps(T)))
    G = F.DenseOperator(A, backend)
    values = []
    for nb in (1, 2):
        _, S = minirsvd(backend, G, T, N, k, plan_for(backend), w, np.random.default_rng(0xBEEF), nbuffers=nb)
        values.append(np.linalg.svd(S, compute_uv=False))
    assert np.allclose(values[0], values[1], rtol=blastol(T, N, k), atol=0)


class CircularConvolution:
    """A matrix-free operator applied one column at a time (test/operators.jl:36-73):
    G[i, j] = kernel[(i - j) mod N], its adjoint the conjugate reversed kernel.
    The column product runs on the host between synchronous copies on the
    current stream, which inside a sweep is the sweep's compute queue."""

    panel_capable = False
    workspace_bytes = 0

    def __init__(self, kernel, backend, conjugated=False):
        self.kernel = np.asarray(kernel)
        self.backend = backend
        self.conjugated = conjugated
        self.dtype = self.kernel.dtype
        self.shape = (len(kernel), len(kernel))

    def adjoint(self):
        return CircularConvolution(self.kernel, self.backend, not self.conjugated)

    def dense(self):
        N = len(self.kernel)
        i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        if self.conjugated:
            return self.kernel[(j - i) % N].conj()
        return self.kernel[(i - j) % N]

    def apply(self, y, x):
        column = self.dense() @ self.backend.to_host(x)
        self.backend.h2d(y, np.asfortranarray(column.astype(self.dtype)))
        return y


def test_mini_rsvd_through_a_matrix_free_operator(backend):
    T = np.complex128
    N, k, w, rank = 120, 24, 7, 6
    rng = np.random.default_rng(0xDEADBEEF)
    # Retained values halve, the rest sit on a floor (test/test_rsvd.jl:52-56);
    # a circulant's eigenvalues are the DFT of its kernel.
    spectrum = np.array([2.0 ** -(m - 1) if m <= rank else 1e-6 for m in range(1, N + 1)])
    G = CircularConvolution(np.fft.ifft(spectrum).astype(T), backend)
    assert F.check_operator(G, backend=backend)
    _, S = minirsvd(backend, G, T, N, k, plan_for(backend), w, rng)
    got = np.linalg.svd(S, compute_uv=False)[:rank]
    assert np.allclose(got, np.linalg.svd(G.dense(), compute_uv=False)[:rank], rtol=1e-6, atol=0)
