"""The disk-tier randomized range finder with power iterations, at a reduced
size: G = the periodic 7-point Helmholtz stencil on a 32^3 grid (N = 32768,
diagonal 6 - k0^2 (1 + 0.05i), k0 = 0.3, neighbours -1), Omega = GhostPanels
(k = 64 in panels of 8, seed 0xdeadbeef), Y = G Omega, then two iterations of
(Z = G^H Y, orthonormalize Z, Y = G Z, orthonormalize Y), every
orthonormalization cholqr2_panelwise with the mixed first pass. Y and Z (16
panels between them) live in a scratch directory behind a host budget of 8
panels, what a two-matrix panel sweep may pin at once (FEED_DEPTH + 2 steps
of one panel per matrix), so every step moves panels through the disk tier. The loop is the
reference demo's range finder (docs/demo/minirsvd.jl:143-154) with G^H and G
alternating; test/test_disk.jl holds the reference's disk-tier tests. The
oracle runs the same chain in FP64 on the CPU."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import plan as P
from oracle import ghost as OG
from oracle import operators as OO
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu

n, k, w = 32, 64, 8
N = n ** 3


def oracle_chain(d, iterations):
    Omega = OG.matrix(np.complex128, N, k, 0xDEADBEEF)
    Y = OO.helmholtz_apply(Omega, (n, n, n), d)
    Y, R = OB.cholqr2_panelwise(Y, w)
    for _ in range(iterations):
        Z = OO.helmholtz_apply(Y, (n, n, n), d, adjoint=True)
        Z, _ = OB.cholqr2_panelwise(Z, w)
        Y = OO.helmholtz_apply(Z, (n, n, n), d)
        Y, R = OB.cholqr2_panelwise(Y, w)
    return Y, R


@pytest.mark.parametrize("nb", (1, 2))
def test_range_finder_through_the_disk_tier(backend, tmp_path, nb):
    print(f"range finder: {n}^3 grid, k={k}, w={w}, nbuffers={nb}", flush=True)
    panel = P.align_up(P.panel_bytes(N, w, np.dtype(np.complex128)), P.SLAB_ALIGN)
    pinned = 2 * (P.FEED_DEPTH + 2)
    plan = F.ResidencyPlan(backend=backend, device_budget=256 << 20, host_budget=pinned * panel + (1 << 20),
                           scratch_dir=str(tmp_path), panel_width=w)
    d = OO.helmholtz_diagonal()
    G = F.Helmholtz3D(n, n, n, d, -1.0, backend=backend)
    Gh = G.adjoint()
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    Z = F.PanelMatrix(N, k, plan=plan, w=w)
    try:
        F.panelmul(Y, G, omega, nbuffers=nb)
        R = F.cholqr2_panelwise(Y, mixed=True, nbuffers=nb)
        for _ in range(2):
            F.panelmul(Z, Gh, Y, nbuffers=nb)
            F.cholqr2_panelwise(Z, mixed=True, nbuffers=nb)
            F.panelmul(Y, G, Z, nbuffers=nb)
            R = F.cholqr2_panelwise(Y, mixed=True, nbuffers=nb)
        assert F.disk_reads(Y) > 0 and F.disk_writes(Y) > 0
        assert F.disk_reads(Z) > 0 and F.disk_writes(Z) > 0
        orth = float(np.linalg.norm(F.gram_panelwise(Y) - np.eye(k)))
        Q = F.to_matrix(Y)
    finally:
        F.free(Y)
        F.free(Z)
    Qo, Ro = oracle_chain(d, 2)
    print(f"||Q^H Q - I||_F = {orth:.2e}, Q off by {np.linalg.norm(Q - Qo) / np.linalg.norm(Qo):.2e}, "
          f"R off by {np.linalg.norm(R - Ro) / np.linalg.norm(Ro):.2e}", flush=True)
    assert orth <= 1e-13
    # The mixed first passes change the rounding of each step, not the
    # subspace: Q and R agree with the FP64 oracle chain to well within the
    # conditioning of the stencil (|G| spans 0.09 to 12).
    assert np.linalg.norm(R - Ro) <= 1e-10 * np.linalg.norm(Ro)
    assert np.linalg.norm(Q - Qo) <= 1e-10 * np.linalg.norm(Qo)
