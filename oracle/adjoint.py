"""Adjoints of cholqr2 and project, restated on the CPU.

TEST INFRASTRUCTURE (see oracle/__init__.py). The reference has no adjoints;
these follow the textbook derivations and are pinned by central finite
differences in tests/test_oracle_adjoint.py.

QR (Y = Q R, Q^H Q = I, R upper with a real positive diagonal), for a real
loss with cotangents Qbar, Rbar:
    M = R Rbar^H - Qbar^H Q,   N = tril(M, -1) + tril(M, -1)^H + Re diag(M),
    Ybar = (Qbar + Q N) R^-H.
project (S = Q^H G Q, G = diag(d) + a fixed part), cotangent Sbar:
    Qbar = G (Q Sbar^H) + G^H (Q Sbar),   dbar_i = Re sum_a conj(Q_ia) (Q Sbar^H)_ia.
"""
import numpy as np
import scipy.linalg


def copyltu_real_diag(M):
    L = np.tril(M, -1)
    return L + L.conj().T + np.diag(np.diag(M).real)


def qr_pullback(Q, R, Qbar, Rbar):
    M = R @ Rbar.conj().T - Qbar.conj().T @ Q
    W = Qbar + Q @ copyltu_real_diag(M)
    # W R^-H = (R^-1 W^H)^H
    return scipy.linalg.solve_triangular(R, W.conj().T, lower=False).conj().T


def project_pullback(Q, apply, apply_adjoint, Sbar):
    Z1 = Q @ Sbar.conj().T
    Z2 = Q @ Sbar
    Qbar = apply(Z1) + apply_adjoint(Z2)
    dbar = np.real(np.sum(Q.conj() * Z1, axis=1))
    return Qbar, dbar
