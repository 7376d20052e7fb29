"""GhostPanels with a host column generator on the HIP backend: the panels
are generated on the producer thread into host blocks and uploaded, as the
reference does (src/ghosts.jl:97-122), and every sweep sees the same matrix
as the host generation alone gives."""
import numpy as np
import pytest

import funicular_hip as F
from oracle import operators as OO
from oracle import panelblas as OB

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("host_eltype", [None, np.complex64])
def test_host_generated_ghost_panels_feed_the_sweeps(backend, host_eltype):
    N, k, w = 20000, 24, 8
    plan = F.ResidencyPlan(backend=backend, device_budget=256 << 20, host_budget=256 << 20, host_eltype=host_eltype)
    O = F.GhostPanels(F.randn_column, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    A = F.to_matrix(O)  # the host generation alone (rounded to the host eltype)
    if host_eltype is not None:
        assert np.array_equal(A, A.astype(host_eltype).astype(np.complex128))
    d, o = 2 + 0.1j, -1.0
    G = F.Tridiagonal1D(N, d, o, backend=backend)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    F.panelmul(Y, G, O)
    # Y is computed in FP64 and kept in the host eltype, so a complex64 host
    # tier rounds it once more.
    tol = 1e-14 if host_eltype is None else 2.0 ** -23
    assert np.allclose(F.to_matrix(Y), OO.tridiag_apply(A, d, o), rtol=tol, atol=tol)
    assert OB.isapprox(F.gram(O), A.conj().T @ A, OB.blastol(np.complex128, N, k))
    # The device stream and the host generator are different streams.
    D = F.to_matrix(F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w))
    assert not np.array_equal(D, A)
