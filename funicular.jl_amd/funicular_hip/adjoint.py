"""Reverse-mode rules for cholqr2 and project, the Python face of the
ChainRulesCore rrules a Julia user would call. The reference has no
derivatives; these follow the QR adjoint (Walter and Lehmann; the complex
form with a real-diagonal R) and the product rule for S = Q^H G Q.

Each rrule runs the forward operation and returns its result with a
pullback. The pullbacks are sweeps over the same row blocks and panels as the
forward pass, on the same kernels (the Gram, the row-block product, the
triangular solve in its conjugate-transpose form), so they stream exactly as
the forward pass does and store no intermediates beyond their outputs.

QR pullback (Y = Q R):  M = R Rbar^H - Qbar^H Q,
                        N = tril(M, -1) + tril(M, -1)^H + Re diag(M),
                        Ybar = (Qbar + Q N) R^-H.
project pullback:       Qbar = G (Q Sbar^H) + G^H (Q Sbar),
                        dbar_i = Re sum_a conj(Q_ia) (Q Sbar^H)_ia
                        for an operator whose diagonal carries parameters d.
"""
import numpy as np

from . import plan as _plan
from .executor import readpanels, rowtraversal, sweep, updatepanels, writepanels
from .panelblas import cholqr2, gram, panelmul, project, rightmul
from .panels import PanelMatrix, assert_same_rows, pooled_like, similar


def copyltu_real_diag(M):
    L = np.tril(M, -1)
    return L + L.conj().T + np.diag(np.diag(M).real)


def rrule_cholqr2(Y, *, nbuffers=None):
    """Runs cholqr2(Y) (Y then holds Q) and returns (R, pullback), where
    pullback(Qbar, Rbar) -> Ybar for a PanelMatrix cotangent Qbar of Q and a
    k x k cotangent Rbar of R (either may be None for zero)."""
    R = cholqr2(Y, nbuffers=nbuffers)
    Q = Y

    def pullback(Qbar=None, Rbar=None):
        plan = Q.plan
        T = Q.dtype
        backend = plan.backend
        k = Q.k
        M = np.zeros((k, k), dtype=T)
        if Rbar is not None:
            M += R @ np.asarray(Rbar).conj().T
        if Qbar is not None:
            assert_same_rows(Qbar, Q)
            M -= gram(Qbar, Q, nbuffers=nbuffers)
        Nm = copyltu_real_diag(M)
        Ybar = PanelMatrix(Q.N, k, plan=plan, w=Q.w, dtype=T)
        Nd, Rd = _plan.checkout_device_buffers(plan, T, (k, k), 2)
        work = None
        try:
            work = _plan.checkout_workspace(plan, backend.trsm_workspace_bytes(k, T))
            backend.h2d(Nd, np.asfortranarray(Nm.astype(T)))
            backend.h2d(Rd, np.asfortranarray(R.astype(T)))

            def step(block, Qb, *rest):
                Yb = rest[-1]
                backend.rightmul_gemm(Yb, Qb, Nd)
                if Qbar is not None:
                    backend.axpby(Yb, rest[0], 1.0, 1.0)
                # The first block folds R^H into the workspace, the rest reuse it.
                if block == 0:
                    backend.rdiv_upper_adjoint(Yb, Rd, work)
                else:
                    backend.rdiv_upper_folded(Yb, work, adjoint=True)
            operands = [readpanels(Q)] + ([readpanels(Qbar)] if Qbar is not None else []) + [writepanels(Ybar)]
            sweep(step, *operands, nbuffers=nbuffers, traversal=rowtraversal(Q))
        finally:
            _plan.checkin_device_buffers(plan, [Nd, Rd, work])
        return Ybar

    return R, pullback


def rrule_project(Q, G, *, nbuffers=None):
    """Runs S = project(Q, G) and returns (S, pullback), where pullback(Sbar)
    -> (Qbar, dbar): Qbar a PanelMatrix, dbar the sensitivity to the real
    diagonal parameters of G (a host vector of length N; the diagonal is the
    operator's `diag` for Tridiagonal1D, the whole diagonal otherwise)."""
    S = project(Q, G, nbuffers=nbuffers)

    def pullback(Sbar):
        Sbar = np.asarray(Sbar, dtype=Q.dtype)
        plan = Q.plan
        backend = plan.backend
        Qbar = similar(Q)
        dbar = backend.alloc_device(np.float64, (Q.N, 1))
        # The two products with Sbar are scratch: under a device-resident plan
        # they come out of (and go back to) the plan's device pool.
        with pooled_like(Q) as Z1, pooled_like(Q) as Z2:
            rightmul(Z1, Q, Sbar.conj().T, nbuffers=nbuffers)
            rightmul(Z2, Q, Sbar, nbuffers=nbuffers)

            def rows(j, Qb, Zb):
                backend.rowdot(dbar, Qb, Zb, accumulate=j > 0)
            sweep(rows, readpanels(Q), readpanels(Z1), nbuffers=nbuffers)
            panelmul(Qbar, G, Z1, nbuffers=nbuffers)
            W = Z1  # reused: Z1 is no longer needed
            panelmul(W, G.adjoint(), Z2, nbuffers=nbuffers)
            sweep(lambda _, Wb, Qb: backend.axpby(Qb, Wb, 1.0, 1.0), readpanels(W), updatepanels(Qbar),
                  nbuffers=nbuffers)
        return Qbar, backend.to_host(dbar).ravel()

    return S, pullback
