"""The operator contract (src/operator.jl:1-183) and the device operators this
backend ships.

An operator G passed to panelmul or project provides `shape`, `dtype`,
`adjoint()` and `apply(Y, X)` for device arrays X (n x c) and Y (m x c).
Three optional traits: `workspace_bytes` (device memory it needs while
applied, default 0), `panel_capable` (apply takes whole panels, default
False: panelmul then applies it one column at a time), `ishermitian_op`
(default False). check_operator holds an operator to all of it.

Shipped operators:
  * DenseOperator  -- a device matrix; G X on the row-block product kernel,
                      G^H X on the Gram kernel.
  * Tridiagonal1D  -- periodic (Gx)_i = d_i x_i + o (x_{i-1} + x_{i+1}).
  * Helmholtz3D    -- periodic 7-point stencil on n1 x n2 x n3.
"""
import numpy as np

from . import _lib
from ._lib import call
from .device_api import ArgumentError, DeviceArray, current_stream, eltype_code


def workspace_bytes(G):
    return int(getattr(G, "workspace_bytes", 0) or 0)


def panel_capable(G):
    return bool(getattr(G, "panel_capable", False))


def ishermitian_op(G):
    return bool(getattr(G, "ishermitian_op", False))


def operator_size(G):
    """src/operator.jl:61-66."""
    dims = getattr(G, "shape", None)
    if not (isinstance(dims, tuple) and len(dims) == 2 and all(isinstance(d, (int, np.integer)) for d in dims)):
        raise ArgumentError(f"G.shape must be a pair of integers, got {dims!r}")
    if not all(d > 0 for d in dims):
        raise ArgumentError(f"operator dimensions must be positive, got {dims}")
    return int(dims[0]), int(dims[1])


def assert_square_operator(G, N):
    rows, cols = operator_size(G)
    if not (rows == cols == N):
        raise ArgumentError(f"this needs a square operator of order {N}, the row count of the panel matrices; "
                            f"the operator is {rows}×{cols}")


def assert_operator_shape(G, m, n):
    rows, cols = operator_size(G)
    if (rows, cols) != (int(m), int(n)):
        raise ArgumentError(f"mapping columns of {n} rows onto columns of {m} rows takes an {m}×{n} operator; "
                            f"this one is {rows}×{cols}")


def _stencil_eltypes(G, Y, X):
    """The stencil kernels read and write the operator's own eltype."""
    if X.dtype != G.dtype or Y.dtype != G.dtype:
        raise ArgumentError(f"a {G.dtype} stencil maps {G.dtype} columns, got {X.dtype} onto {Y.dtype}. Build the "
                            f"operator with dtype={X.dtype} to apply it to them")


def apply_operator(dst, G, src):
    """src/operator.jl:80."""
    if panel_capable(G):
        G.apply(dst, src)
    else:
        columnwise_apply(dst, G, src)
    return dst


def columnwise_apply(dst, G, src):
    """src/operator.jl:82-87."""
    for c in range(src.cols):
        G.apply(dst.view(None, (c, c + 1)), src.view(None, (c, c + 1)))
    return dst


class DenseOperator:
    """A dense device matrix as an operator (a plain matrix satisfies the
    reference's contract and is panel capable, src/operator.jl:25)."""

    panel_capable = True
    workspace_bytes = 0

    def __init__(self, A, backend, _adjoint_of=None):
        self.backend = backend
        if _adjoint_of is not None:
            self.A = _adjoint_of
            self._adj = True
            self.shape = (self.A.cols, self.A.rows)
        else:
            self.A = A if isinstance(A, DeviceArray) else backend.todevice(np.asfortranarray(A))
            self._adj = False
            self.shape = self.A.shape
        self.dtype = self.A.dtype
        self._herm = None

    @property
    def ishermitian_op(self):
        if self._herm is None:
            if self.shape[0] != self.shape[1]:
                self._herm = False
            else:
                H = self.backend.to_host(self.A)
                self._herm = bool(np.array_equal(H, H.conj().T))
        return self._herm

    def adjoint(self):
        if self._adj:
            return DenseOperator(self.A, self.backend)
        return DenseOperator(None, self.backend, _adjoint_of=self.A)

    def apply(self, Y, X):
        if self._adj:
            # G^H X as the Gram kernel's A^H B; its workspace is small (m = rows of G).
            self.backend.gram_accumulate(Y, self.A, X, 1.0, 0.0)
        else:
            self.backend.rightmul_gemm(Y, self.A, X)
        return Y


class Tridiagonal1D:
    """Periodic tridiagonal stencil (Gx)_i = (d + diag_i) x_i + o (x_{i-1} + x_{i+1}),
    indices mod N. The 1-D sketch chain uses d = 2+0.1i, o = -1; the adjoint
    workload adds the real parameters diag_i = 1 + 0.5 sin(2 pi i / N)."""

    panel_capable = True
    workspace_bytes = 0

    def __init__(self, N, d, o, diag=None, backend=None, dtype=np.complex128, _adjoint=False):
        if backend is None:
            from .device_api import hip_backend
            backend = hip_backend()
        self.backend = backend
        self.N = int(N)
        self.d = complex(d)
        self.o = complex(o)
        self.dtype = np.dtype(dtype)
        self.shape = (self.N, self.N)
        self._adjoint = _adjoint
        if diag is None:
            self.diag = None
        elif isinstance(diag, DeviceArray):
            self.diag = diag
        else:
            diag = np.ascontiguousarray(np.asarray(diag, dtype=np.float64))
            if diag.shape != (self.N,):
                raise ArgumentError(f"diag must have {self.N} entries")
            self.diag = backend.todevice(diag.reshape(-1, 1))

    @property
    def ishermitian_op(self):
        return self.d.imag == 0 and self.o.imag == 0

    def adjoint(self):
        return Tridiagonal1D(self.N, self.d, self.o, self.diag, self.backend, self.dtype, not self._adjoint)

    def apply(self, Y, X):
        if X.rows != self.N or Y.rows != self.N or X.cols != Y.cols:
            raise ArgumentError(f"a {self.N}×{self.N} stencil cannot map {X.shape} onto {Y.shape}")
        _stencil_eltypes(self, Y, X)
        call("fhip_stencil_tridiag", Y.ptr, Y.ld, X.ptr, X.ld, self.N, X.cols, self.d.real, self.d.imag,
             self.o.real, self.o.imag, None if self.diag is None else self.diag.ptr, 1 if self._adjoint else 0,
             eltype_code(self.dtype), current_stream())
        return Y

    def dense(self):
        """The operator as a dense host matrix (for small checks)."""
        N = self.N
        d = np.full(N, self.d, dtype=np.complex128)
        if self.diag is not None:
            d = d + self.backend.to_host(self.diag).ravel()
        G = np.diag(d)
        for i in range(N):
            G[i, (i - 1) % N] += self.o
            G[i, (i + 1) % N] += self.o
        return G.conj().T if self._adjoint else G


class Helmholtz3D:
    """Periodic 7-point stencil on n1 x n2 x n3 (linear index i + n1 (j + n2 l)):
    (Gx) = d x + o * (sum of the six neighbours). Configurations 2 and 5 use
    d = 6 - k0^2 (1 + 0.05i) with k0 = 0.3, o = -1."""

    panel_capable = True
    workspace_bytes = 0

    def __init__(self, n1, n2, n3, d, o=-1.0, backend=None, dtype=np.complex128, _adjoint=False):
        if backend is None:
            from .device_api import hip_backend
            backend = hip_backend()
        self.backend = backend
        self.n = (int(n1), int(n2), int(n3))
        self.N = self.n[0] * self.n[1] * self.n[2]
        self.d = complex(d)
        self.o = complex(o)
        self.dtype = np.dtype(dtype)
        self.shape = (self.N, self.N)
        self._adjoint = _adjoint

    @property
    def ishermitian_op(self):
        return self.d.imag == 0 and self.o.imag == 0

    def adjoint(self):
        n1, n2, n3 = self.n
        return Helmholtz3D(n1, n2, n3, self.d, self.o, self.backend, self.dtype, not self._adjoint)

    def apply(self, Y, X):
        if X.rows != self.N or Y.rows != self.N or X.cols != Y.cols:
            raise ArgumentError(f"a {self.N}×{self.N} stencil cannot map {X.shape} onto {Y.shape}")
        _stencil_eltypes(self, Y, X)
        n1, n2, n3 = self.n
        call("fhip_stencil_helmholtz3d", Y.ptr, Y.ld, X.ptr, X.ld, n1, n2, n3, X.cols, self.d.real, self.d.imag,
             self.o.real, self.o.imag, 1 if self._adjoint else 0, eltype_code(self.dtype), current_stream())
        return Y


def randprobe(T, m, n, rng):
    """src/operator.jl:176-177."""
    T = np.dtype(T)
    if np.issubdtype(T, np.complexfloating):
        return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(T)
    return rng.standard_normal((m, n)).astype(T)


def check_operator(G, n=3, backend=None, rtol=None, seed=None):
    """Holds G to the contract panelmul and project expect, with n random
    probes; raises ArgumentError naming the first failure (src/operator.jl:89-174)."""
    if backend is None:
        backend = getattr(G, "backend", None)
        if backend is None:
            from .device_api import hip_backend
            backend = hip_backend()
    probes = int(n)
    if probes < 1:
        raise ArgumentError(f"check_operator needs one probe or more; n = {n}")
    rows, cols = operator_size(G)
    try:
        T = np.dtype(G.dtype)
    except Exception:
        raise ArgumentError(f"G.dtype must be a concrete number type, got {getattr(G, 'dtype', None)!r}")
    if not backend.supports_eltype(T):
        raise ArgumentError(f"backend {type(backend).__name__} cannot compute in the operator's eltype {T}")
    if workspace_bytes(G) < 0:
        raise ArgumentError(f"the operator reports a negative workspace ({workspace_bytes(G)} bytes)")
    adj = G.adjoint()
    adjdims = operator_size(adj)
    if adjdims != (cols, rows):
        raise ArgumentError(f"G.adjoint() has shape {adjdims[0]}×{adjdims[1]}, not the transposed shape "
                            f"{cols}×{rows} of G")
    real = np.finfo(T).eps
    tol = 64 * real * np.sqrt(max(rows, cols)) if rtol is None else rtol
    rng = np.random.default_rng(seed)
    Xh = randprobe(T, cols, probes, rng)
    Yh = randprobe(T, rows, probes, rng)
    X, Y = backend.todevice(Xh), backend.todevice(Yh)
    GX = backend.alloc_device(T, (rows, probes))
    AY = backend.alloc_device(T, (cols, probes))
    backend.fill(GX, 0)
    backend.fill(AY, 0)
    columnwise_apply(GX, G, X)
    columnwise_apply(AY, adj, Y)
    GXh, AYh = backend.to_host(GX), backend.to_host(AY)
    for c in range(probes):
        left = np.vdot(Yh[:, c], GXh[:, c])
        right = np.vdot(AYh[:, c], Xh[:, c])
        scale = max(abs(left), abs(right), 1.0)
        if not abs(left - right) <= max(tol * max(abs(left), abs(right)), tol * scale):
            raise ArgumentError(
                f"probe {c}: <y, G x> = {left} while <G^H y, x> = {right} (relative gap "
                f"{abs(left - right) / max(abs(left), abs(right))}, allowed {tol}), so G.adjoint() is not the "
                "adjoint of G")
    if panel_capable(G):
        panelled = backend.alloc_device(T, (rows, probes))
        backend.fill(panelled, 0)
        try:
            G.apply(panelled, X)
        except Exception as err:
            raise ArgumentError(f"G sets panel_capable, yet applying it to a {cols}×{probes} block raised: {err}")
        P = backend.to_host(panelled)
        if not np.allclose(P, GXh, rtol=tol, atol=tol):
            raise ArgumentError(f"G sets panel_capable, yet its result on a {cols}×{probes} block differs from "
                                f"its column-by-column result by more than {tol} (relative)")
    if ishermitian_op(G):
        if rows != cols:
            raise ArgumentError(f"G sets ishermitian_op, yet it is {rows}×{cols}, not square")
        GY = backend.alloc_device(T, (rows, probes))
        backend.fill(GY, 0)
        columnwise_apply(GY, G, Y)
        GYh = backend.to_host(GY)
        for c in range(probes):
            left = np.vdot(Yh[:, c], GXh[:, c])
            right = np.vdot(GYh[:, c], Xh[:, c])
            scale = max(abs(left), abs(right), 1.0)
            if not abs(left - right) <= tol * scale:
                raise ArgumentError(f"G sets ishermitian_op, yet probe {c} gives <y, G x> = {left} and "
                                    f"<G y, x> = {right}, apart by more than {tol}")
    return True
