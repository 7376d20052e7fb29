"""The device contract of Funicular (src/device_api.jl:12-138) on HIP.

`HIPBackend` is what `hip_backend()` returns: device buffers from hipMalloc,
page-locked host slabs, non-blocking HIP streams as queues, HIP events as
queue ordering, and the panel kernels of libfunicular_hip.so in place of the
vendor BLAS calls of src/kernels.jl. The behavioural model is the CUDA
extension (ext/FunicularCUDAExt.jl): everything is asynchronous on an explicit
stream, and the only host waits are the ones the caller asks for.

Where CUDA.jl kept a task-local current stream (ext/FunicularCUDAExt.jl:83-89),
this module keeps a thread-local one: `submit` makes the queue's stream
current while the panel function runs, and every kernel wrapper launches on
the current stream.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import call

SLAB_ALIGN = 64  # src/device_api.jl:140

_ELTYPE_CODES = {
    np.dtype(np.float32): _lib.FHIP_F32,
    np.dtype(np.float64): _lib.FHIP_F64,
    np.dtype(np.complex64): _lib.FHIP_C32,
    np.dtype(np.complex128): _lib.FHIP_C64,
}


def eltype_code(T):
    try:
        return _ELTYPE_CODES[np.dtype(T)]
    except KeyError:
        raise ArgumentError(f"no HIP kernel works in eltype {np.dtype(T)}; use float32, float64, complex64 or complex128")


class ArgumentError(ValueError):
    """Julia's ArgumentError: a contract or budget violation."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch: shapes that do not line up."""


def human_readable_bytes(n):
    """src/device_api.jl:301-311."""
    n = int(n)
    if n < 1024:
        return f"{n} B"
    units = ("KiB", "MiB", "GiB", "TiB")
    x = n / 1024
    i = 0
    while x >= 1024 and i < len(units) - 1:
        x /= 1024
        i += 1
    return f"{round(x, 2)} {units[i]}"


# ---------------------------------------------------------------- buffers --

class DeviceBuffer:
    """One hipMalloc allocation, freed when the last array over it goes."""

    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        call("fhip_malloc", ctypes.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            _lib.load().fhip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A column-major device matrix or a view of one.

    `ptr` is the address of element (0, 0), `ld` the distance between columns
    in elements, `owner` the allocation that keeps the memory alive. A view
    (`view(rows, cols)`) shares the owner and the leading dimension, which is
    what a Julia SubArray of a CuArray is to the CUDA extension's copies and
    kernels (ext/FunicularCUDAExt.jl:157-182).
    """

    __slots__ = ("ptr", "shape", "ld", "dtype", "owner")

    def __init__(self, ptr, shape, ld, dtype, owner):
        self.ptr = int(ptr)
        self.shape = (int(shape[0]), int(shape[1]))
        self.ld = int(ld)
        self.dtype = np.dtype(dtype)
        self.owner = owner

    @property
    def rows(self):
        return self.shape[0]

    @property
    def cols(self):
        return self.shape[1]

    @property
    def size(self):
        return self.shape[0] * self.shape[1]

    # In-place elementwise updates on the current stream, the broadcasts a
    # foreachpanel closure applies to its buffer in Julia (panel .*= a,
    # panel .+= X, panel .= a .* panel .+ b, panel .= v; src/executor.jl:241-267).
    def __imul__(self, alpha):
        _array_backend().scale(self, alpha)
        return self

    def __iadd__(self, other):
        if isinstance(other, DeviceArray):
            _array_backend().axpby(self, other, 1.0, 1.0)
        else:
            _array_backend().affine(self, 1.0, other)
        return self

    def __isub__(self, other):
        if isinstance(other, DeviceArray):
            _array_backend().axpby(self, other, -1.0, 1.0)
        else:
            _array_backend().affine(self, 1.0, -other)
        return self

    def affine_(self, alpha, beta):
        """self <- alpha self + beta (Julia's self .= alpha .* self .+ beta)."""
        _array_backend().affine(self, alpha, beta)
        return self

    def fill_(self, value):
        """self <- value everywhere (Julia's self .= value)."""
        _array_backend().fill(self, value)
        return self

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    def __len__(self):
        return self.shape[0]

    def view(self, rows=None, cols=None):
        """Row range and column range, 0-based half-open (start, stop) pairs or slices."""
        r0, r1 = _span(rows, self.shape[0])
        c0, c1 = _span(cols, self.shape[1])
        ptr = self.ptr + (r0 + c0 * self.ld) * self.dtype.itemsize
        return DeviceArray(ptr, (r1 - r0, c1 - c0), self.ld, self.dtype, self.owner)

    def is_contiguous(self):
        return self.ld == self.shape[0] or self.shape[1] <= 1

    def __repr__(self):
        return f"DeviceArray({self.shape[0]}x{self.shape[1]} {self.dtype}, ld={self.ld})"


def _span(r, n):
    if r is None:
        return 0, n
    if isinstance(r, slice):
        start, stop, step = r.indices(n)
        if step != 1:
            raise ArgumentError("device views take unit-stride ranges only")
        return start, max(start, stop)
    start, stop = int(r[0]), int(r[1])
    if not (0 <= start <= stop <= n):
        raise IndexError(f"range {start}:{stop} outside 0:{n}")
    return start, stop


class HostSlab:
    """One page-locked host allocation (src/device_api.jl:140-163)."""

    __slots__ = ("ptr", "nbytes", "_raw", "__weakref__")

    def __init__(self, nbytes):
        raw = ctypes.c_void_p()
        call("fhip_host_alloc", ctypes.byref(raw), int(nbytes) + SLAB_ALIGN)
        self._raw = raw.value
        pad = (-self._raw) % SLAB_ALIGN
        self.ptr = self._raw + pad
        self.nbytes = int(nbytes)

    def __len__(self):
        return self.nbytes

    def free(self):
        if self._raw:
            _lib.load().fhip_host_free(self._raw)
            self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def slab_matrix(slab, T, dims, offset):
    """A Fortran-ordered numpy view of `dims` elements of T at `offset` bytes
    into the slab (src/device_api.jl:282-287)."""
    T = np.dtype(T)
    rows, cols = int(dims[0]), int(dims[1])
    nbytes = rows * cols * T.itemsize
    if offset < 0:
        raise ArgumentError(f"slab offset must be nonnegative, got {offset}")
    if offset + nbytes > slab.nbytes:
        raise ArgumentError(
            f"{human_readable_bytes(nbytes)} starting {offset} bytes in runs past the end of a "
            f"{human_readable_bytes(slab.nbytes)} host slab")
    if nbytes == 0:
        return np.zeros((rows, cols), dtype=T, order="F")
    raw = np.asarray(_SlabRegion(slab, offset, nbytes))
    return raw.view(T).reshape((rows, cols), order="F")


class _SlabRegion:
    """Exposes bytes of a slab to numpy; the arrays built over it keep this
    object, and through it the slab, alive."""

    def __init__(self, slab, offset, nbytes):
        self.slab = slab
        self.__array_interface__ = {
            "shape": (int(nbytes),), "typestr": "|u1", "version": 3,
            "data": (slab.ptr + int(offset), False)}


def host_region(A):
    """(pointer, ld in elements) of a 2-D host array with contiguous columns
    (ext/FunicularCUDAExt.jl:157-163)."""
    if A.ndim == 1:
        A = A.reshape(-1, 1)
    if A.ndim != 2:
        raise ArgumentError(f"copies move matrices (panels and row blocks); got an array of {A.ndim} dimensions")
    item = A.dtype.itemsize
    if A.shape[0] > 1 and A.strides[0] != item:
        raise ArgumentError(
            f"copies need each column contiguous in memory, but consecutive rows of this array are "
            f"{A.strides[0] // item} elements apart")
    ld = A.strides[1] // item if A.shape[1] > 1 else max(A.shape[0], 1)
    return A.ctypes.data, max(ld, A.shape[0])


# ------------------------------------------------------------ queues, events --

_current = threading.local()


def current_stream():
    """The stream the calling thread's kernels launch on (None = null stream)."""
    return getattr(_current, "stream", None)


class Queue:
    """A HIP stream. Streams are recycled through the backend's pool."""

    __slots__ = ("stream", "_pool", "__weakref__")

    def __init__(self, stream, pool):
        self.stream = stream
        self._pool = pool

    def __del__(self):
        pool = self._pool
        if pool is not None and self.stream:
            pool.append(self.stream)
            self.stream = None


class Event:
    """A HIP event recorded on a queue (timing disabled)."""

    __slots__ = ("handle",)

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        h = self.handle
        if h:
            self.handle = None
            try:
                _lib.load().fhip_event_destroy(h)
            except Exception:
                pass


class DeviceBackend:
    """Abstract device backend (src/device_api.jl:5)."""

    def supports_eltype(self, T):
        return True

    def transfers_are_real(self):
        return True

    def make_sweep_queues(self):
        return (self.make_queue(), self.make_queue(), self.make_queue())


class HIPBackend(DeviceBackend):
    """Funicular's device backend on an AMD Instinct GPU (gfx950).

    Panels live in hipMalloc buffers, host slabs are page-locked, queues are
    non-blocking streams, and ordering between queues is HIP events waited on
    by the device. The backend holds no data: staging buffers belong to the
    ResidencyPlan, streams to whichever sweep took them.
    """

    def __init__(self, device=0):
        lib = _lib.load()
        n = ctypes.c_int()
        call("fhip_device_count", ctypes.byref(n))
        if n.value < 1:
            raise RuntimeError("hip_backend: no HIP device is visible")
        if not (0 <= device < n.value):
            raise ArgumentError(f"device {device} is not among the {n.value} visible")
        call("fhip_set_device", int(device))
        self.device = device
        self._stream_pool = []
        buf = ctypes.create_string_buffer(128)
        call("fhip_device_arch", buf, 128)
        self.arch = buf.value.decode()
        self._lib = lib

    def __repr__(self):
        return f"HIPBackend(device {self.device}, {self.arch})"

    # -- memory --
    def devicetype(self, T):
        return DeviceArray

    def alloc_device(self, T, dims):
        T = np.dtype(T)
        if len(dims) == 1:
            dims = (dims[0], 1)
        rows, cols = int(dims[0]), int(dims[1])
        buf = DeviceBuffer(max(rows * cols * T.itemsize, 1))
        return DeviceArray(buf.ptr, (rows, cols), max(rows, 1), T, buf)

    def alloc_host_slab(self, nbytes):
        return HostSlab(nbytes)

    def supports_eltype(self, T):
        return np.dtype(T) in _ELTYPE_CODES

    def transfers_are_real(self):
        return True

    # -- queues and events --
    def make_queue(self):
        if self._stream_pool:
            return Queue(self._stream_pool.pop(), self._stream_pool)
        s = ctypes.c_void_p()
        call("fhip_stream_create", ctypes.byref(s))
        return Queue(s.value, self._stream_pool)

    def record_event(self, queue):
        e = ctypes.c_void_p()
        call("fhip_event_create", ctypes.byref(e), 1)
        ev = Event(e.value)
        call("fhip_event_record", ev.handle, _stream_of(queue))
        return ev

    def wait_event(self, queue, event):
        if event is None:
            return None
        call("fhip_stream_wait_event", _stream_of(queue), event.handle)
        return None

    def submit(self, queue, work):
        """Runs work() now, on the calling thread, with the queue's stream
        current, and returns the event its launches finish on
        (ext/FunicularCUDAExt.jl:83-89)."""
        previous = getattr(_current, "stream", None)
        _current.stream = _stream_of(queue)
        try:
            work()
        finally:
            _current.stream = previous
        return self.record_event(queue)

    def sync_queue(self, queue):
        call("fhip_stream_sync", _stream_of(queue))

    def sync_event(self, event):
        if event is not None:
            call("fhip_event_sync", event.handle)

    def sync_all(self):
        call("fhip_device_sync")

    # -- copies --
    def h2d(self, dst, src, queue=None):
        """Copies host `src` into device `dst` (src/device_api.jl:37)."""
        src = np.asarray(src)
        if src.ndim == 1:
            src = src.reshape(-1, 1)
        if tuple(dst.shape) != tuple(src.shape):
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        if dst.dtype != src.dtype:
            return self._converting_copy(dst, src, queue, to_device=True)
        return self._copy(dst.ptr, dst.ld, src, _lib.FHIP_H2D, queue, host_is_src=True, dev=dst)

    def d2h(self, dst, src, queue=None):
        """Copies device `src` into host `dst` (src/device_api.jl:45)."""
        if dst.ndim == 1:
            dst = dst.reshape(-1, 1)
        if tuple(dst.shape) != tuple(src.shape):
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        if dst.dtype != src.dtype:
            return self._converting_copy(src, dst, queue, to_device=False)
        return self._copy(src.ptr, src.ld, dst, _lib.FHIP_D2H, queue, host_is_src=False, dev=src)

    def d2d(self, dst, src, queue=None):
        if tuple(dst.shape) != tuple(src.shape):
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        if dst.dtype != src.dtype:
            self.convert_copy(dst, src, stream=_stream_of(queue) if queue is not None else current_stream())
            return self._finish(queue)
        item = dst.dtype.itemsize
        stream = _stream_of(queue) if queue is not None else current_stream()
        call("fhip_memcpy2d_async", dst.ptr, dst.ld * item, src.ptr, src.ld * item,
             dst.rows * item, dst.cols, _lib.FHIP_D2D, stream)
        return self._finish(queue)

    def _copy(self, dev_ptr, dev_ld, host, kind, queue, host_is_src, dev):
        if host.size == 0:
            return self._finish(queue)
        hptr, hld = host_region(host)
        item = host.dtype.itemsize
        rows, cols = host.shape if host.ndim == 2 else (host.shape[0], 1)
        stream = _stream_of(queue) if queue is not None else current_stream()
        if host_is_src:
            call("fhip_memcpy2d_async", dev_ptr, dev_ld * item, hptr, hld * item, rows * item, cols, kind, stream)
        else:
            call("fhip_memcpy2d_async", hptr, hld * item, dev_ptr, dev_ld * item, rows * item, cols, kind, stream)
        return self._finish(queue)

    def _finish(self, queue):
        # A copy without a queue is synchronous, on the calling thread's
        # current stream: inside submit that is the queue's, so the copy is
        # ordered after the work already queued there (the task-local stream
        # of ext/FunicularCUDAExt.jl:83-89); outside, the null stream.
        if queue is None:
            call("fhip_stream_sync", current_stream())
            return None
        return self.record_event(queue)

    def _converting_copy(self, dev, host, queue, to_device):
        """A bare copy between eltypes outside a sweep (ext/FunicularCUDAExt.jl:195-212):
        stage in the host eltype on the device, convert there, and wait."""
        staging = self.alloc_device(host.dtype, host.shape)
        stream = _stream_of(queue) if queue is not None else current_stream()
        if to_device:
            self._copy(staging.ptr, staging.ld, host, _lib.FHIP_H2D, None, True, staging)
            self.convert_copy(dev, staging, stream=stream)
            call("fhip_stream_sync", stream)
        else:
            self.convert_copy(staging, dev, stream=stream)
            call("fhip_stream_sync", stream)
            self._copy(staging.ptr, staging.ld, host, _lib.FHIP_D2H, None, False, staging)
        return None

    def to_host(self, A):
        """A fresh host copy of a device array (synchronous)."""
        out = np.empty(A.shape, dtype=A.dtype, order="F")
        self.d2h(out, A)
        return out

    def todevice(self, A):
        """src/operator.jl:179-183."""
        A = np.asfortranarray(A)
        if A.ndim == 1:
            A = A.reshape(-1, 1)
        dev = self.alloc_device(A.dtype, A.shape)
        self.h2d(dev, A)
        return dev

    # -- kernels (the overridable ones of src/kernels.jl) --
    def gram_workspace_bytes(self, m, ka, kb, T, hermitian=True, compute=None):
        n = ctypes.c_size_t()
        code = eltype_code(T)
        ccode = code if compute is None else eltype_code(compute)
        call("fhip_gram_workspace_bytes", ctypes.byref(n), int(m), int(ka), int(kb), code, ccode,
             1 if hermitian else 0)
        return n.value

    def trsm_workspace_bytes(self, k, T):
        n = ctypes.c_size_t()
        call("fhip_trsm_workspace_bytes", ctypes.byref(n), int(k), eltype_code(T))
        return n.value

    def gram_accumulate(self, C, A, B, alpha, beta, work=None, compute=None):
        """C <- alpha A^H B + beta C (src/kernels.jl:40)."""
        if not (A.rows == B.rows and C.shape == (A.cols, B.cols)):
            raise DimensionMismatch(f"cannot accumulate the {A.shape}' times {B.shape} product into a {C.shape} matrix")
        if not (A.dtype == B.dtype == C.dtype):
            raise ArgumentError("gram_accumulate needs one eltype for A, B and C")
        if C.size == 0:
            return C
        herm = A.ptr == B.ptr and A.ld == B.ld and A.cols == B.cols
        need = self.gram_workspace_bytes(A.rows, A.cols, B.cols, A.dtype, herm, compute)
        work_ptr, work_bytes = _workspace(self, work, need)
        code = eltype_code(A.dtype)
        ccode = code if compute is None else eltype_code(compute)
        alpha = complex(alpha)
        beta = complex(beta)
        call("fhip_gram", C.ptr, C.ld, A.ptr, A.ld, B.ptr, B.ld, A.rows, A.cols, B.cols,
             alpha.real, alpha.imag, beta.real, beta.imag, code, ccode, work_ptr, work_bytes,
             current_stream())
        return C

    def rightmul_gemm(self, dst, A, C):
        """dst <- A C (src/kernels.jl:73)."""
        if not (A.cols == C.rows and dst.shape == (A.rows, C.cols)):
            raise DimensionMismatch(f"cannot multiply a {A.shape} block by a {C.shape} factor into a {dst.shape} matrix")
        if dst.size == 0:
            return dst
        call("fhip_gemm_nn", dst.ptr, dst.ld, A.ptr, A.ld, C.ptr, C.ld, A.rows, A.cols, C.cols,
             eltype_code(dst.dtype), current_stream())
        return dst

    def rightmul_gemm_scaled(self, dst, A, C, alpha, beta):
        """dst <- alpha A C + beta dst (beta = 0 does not read dst)."""
        if not (A.cols == C.rows and dst.shape == (A.rows, C.cols)):
            raise DimensionMismatch(f"cannot multiply a {A.shape} block by a {C.shape} factor into a {dst.shape} matrix")
        if not (A.dtype == C.dtype == dst.dtype):
            raise ArgumentError("rightmul_gemm_scaled needs one eltype for A, C and dst")
        if dst.size == 0:
            return dst
        a, b = complex(alpha), complex(beta)
        call("fhip_gemm_nn_scaled", dst.ptr, dst.ld, A.ptr, A.ld, C.ptr, C.ld, A.rows, A.cols, C.cols,
             a.real, a.imag, b.real, b.imag, eltype_code(dst.dtype), current_stream())
        return dst

    def rdiv_upper(self, Y, R, work=None, compute=None):
        """Y <- Y R^-1 for upper triangular R (src/kernels.jl:108)."""
        if not (R.rows == R.cols == Y.cols):
            raise DimensionMismatch(f"a {Y.shape} block cannot be divided by a {R.shape} triangular factor")
        if Y.size == 0:
            return Y
        code = eltype_code(Y.dtype)
        ccode = code if compute is None else eltype_code(compute)
        need = self.trsm_workspace_bytes(R.rows, Y.dtype if compute is None else compute)
        work_ptr, work_bytes = _workspace(self, work, need)
        call("fhip_trsm_runn", Y.ptr, Y.ld, Y.rows, Y.cols, R.ptr, R.ld, code, ccode,
             work_ptr, work_bytes, current_stream())
        return Y

    def potrf_upper(self, R, G, info):
        """R <- the upper Cholesky factor of Hermitian G (upper triangle read),
        on the current stream; `info` is a 1 x 1 int32 device matrix that
        becomes 0 or the 1-based column of the first bad pivot. Returns
        without waiting; read `info` after the stream."""
        k = G.rows
        if not (G.cols == k and R.shape == (k, k)):
            raise DimensionMismatch(f"cannot factor a {G.shape} matrix into a {R.shape} one")
        if info.dtype != np.int32 or info.size < 1:
            raise ArgumentError("potrf_upper needs an int32 device scalar for info")
        call("fhip_potrf_upper", R.ptr, R.ld, eltype_code(R.dtype), G.ptr, G.ld, eltype_code(G.dtype), k, info.ptr,
             current_stream())
        return R

    def rdiv_upper_folded(self, Y, work, compute=None, adjoint=False):
        """The solve of the last rdiv_upper (or rdiv_upper_adjoint) on another
        block, reusing the folded factor it left in `work`."""
        if Y.size == 0:
            return Y
        if work is None:
            raise ArgumentError("rdiv_upper_folded needs the workspace a previous solve folded the factor into")
        code = eltype_code(Y.dtype)
        ccode = code if compute is None else eltype_code(compute)
        need = self.trsm_workspace_bytes(Y.cols, Y.dtype if compute is None else compute)
        work_ptr, work_bytes = _workspace(self, work, need)
        call("fhip_trsm_apply", Y.ptr, Y.ld, Y.rows, Y.cols, code, ccode, 1 if adjoint else 0, work_ptr, work_bytes,
             current_stream())
        return Y

    def rdiv_upper_adjoint(self, Y, R, work=None):
        """Y <- Y R^-H for upper triangular R (the solve of the QR adjoint)."""
        if not (R.rows == R.cols == Y.cols):
            raise DimensionMismatch(f"a {Y.shape} block cannot be divided by a {R.shape} triangular factor")
        if Y.size == 0:
            return Y
        need = self.trsm_workspace_bytes(R.rows, Y.dtype)
        work_ptr, work_bytes = _workspace(self, work, need)
        call("fhip_trsm_runc", Y.ptr, Y.ld, Y.rows, Y.cols, R.ptr, R.ld, eltype_code(Y.dtype), work_ptr,
             work_bytes, current_stream())
        return Y

    def convert_copy(self, dst, src, stream=None):
        """dst <- convert(eltype(dst), src) (src/kernels.jl:14-20)."""
        if tuple(dst.shape) != tuple(src.shape):
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        if dst.size == 0:
            return dst
        s = current_stream() if stream is None else stream
        call("fhip_convert_copy", dst.ptr, eltype_code(dst.dtype), dst.ld, src.ptr, eltype_code(src.dtype),
             src.ld, dst.rows, dst.cols, s)
        return dst

    def axpby(self, Y, X, a, b):
        """Y <- a X + b Y."""
        if tuple(Y.shape) != tuple(X.shape) or Y.dtype != X.dtype:
            raise DimensionMismatch(f"axpby between a {Y.shape} {Y.dtype} and a {X.shape} {X.dtype} matrix")
        a, b = complex(a), complex(b)
        call("fhip_axpby", Y.ptr, Y.ld, X.ptr, X.ld, Y.rows, Y.cols, a.real, a.imag, b.real, b.imag,
             eltype_code(Y.dtype), current_stream())
        return Y

    def scale(self, Y, alpha):
        """Y <- Y alpha entry by entry, as rmul! computes it: a real scalar
        scales both parts, a complex one is the schoolbook product."""
        real_scalar = not np.iscomplexobj(alpha)
        a = complex(alpha)
        if not np.issubdtype(Y.dtype, np.complexfloating) and a.imag != 0:
            raise ArgumentError(f"cannot scale a real {Y.dtype} matrix by the complex number {alpha}")
        code = eltype_code(Y.dtype)
        for c0 in range(0, Y.cols, _MAX_GRID_COLS):
            part = Y.view(None, (c0, min(Y.cols, c0 + _MAX_GRID_COLS)))
            call("fhip_scale", part.ptr, part.ld, part.rows, part.cols, a.real, a.imag, 1 if real_scalar else 0, code,
                 current_stream())
        return Y

    def affine(self, Y, alpha, beta):
        """Y <- alpha Y + beta entry by entry: the product as scale forms it,
        then beta added, each part rounded once per operation (Julia's
        Y .= alpha .* Y .+ beta)."""
        real_scalar = not np.iscomplexobj(alpha)
        a, b = complex(alpha), complex(beta)
        if not np.issubdtype(Y.dtype, np.complexfloating) and (a.imag != 0 or b.imag != 0):
            raise ArgumentError(f"cannot apply the complex coefficients {alpha}, {beta} to a real {Y.dtype} matrix")
        code = eltype_code(Y.dtype)
        for c0 in range(0, Y.cols, _MAX_GRID_COLS):
            part = Y.view(None, (c0, min(Y.cols, c0 + _MAX_GRID_COLS)))
            call("fhip_affine", part.ptr, part.ld, part.rows, part.cols, a.real, a.imag, 1 if real_scalar else 0,
                 b.real, b.imag, code, current_stream())
        return Y

    def dot_workspace_bytes(self, rows, cols):
        n = ctypes.c_size_t()
        call("fhip_dot_workspace_bytes", ctypes.byref(n), int(rows), int(min(cols, _MAX_GRID_COLS)))
        return n.value

    def dot_accumulate(self, out, X, Y=None, accumulate=False, work=None):
        """out (a 1 x 1 complex128 device matrix) <- (accumulate ? out : 0) +
        sum conj(X) .* Y, or sum |X|^2 when Y is None, summed in FP64 in an
        order fixed by the shape."""
        if Y is not None and (tuple(X.shape) != tuple(Y.shape) or X.dtype != Y.dtype):
            raise DimensionMismatch(f"dot of a {X.shape} {X.dtype} and a {Y.shape} {Y.dtype} matrix")
        if out.dtype != np.complex128 or out.size < 1:
            raise ArgumentError("dot_accumulate needs a complex128 device accumulator")
        work_ptr, work_bytes = _workspace(self, work, self.dot_workspace_bytes(X.rows, X.cols))
        code = eltype_code(X.dtype)
        mode = 1 if Y is None else 0
        cols = max(X.cols, 1)
        for c0 in range(0, cols, _MAX_GRID_COLS):
            c1 = min(X.cols, c0 + _MAX_GRID_COLS)
            xp = X.view(None, (c0, c1))
            yp = xp if Y is None else Y.view(None, (c0, c1))
            call("fhip_dot", out.ptr, xp.ptr, xp.ld, yp.ptr, yp.ld, xp.rows, xp.cols, code, mode,
                 1 if (accumulate or c0 > 0) else 0, work_ptr, work_bytes, current_stream())
        return out

    def rowdot(self, out, A, B, accumulate=False):
        """out[i] (+)= Re sum_j conj(A[i, j]) B[i, j]; out is a float64 device vector."""
        if tuple(A.shape) != tuple(B.shape) or A.dtype != B.dtype or out.rows < A.rows:
            raise DimensionMismatch(f"rowdot of {A.shape} and {B.shape} into {out.shape}")
        call("fhip_rowdot", out.ptr, A.ptr, A.ld, B.ptr, B.ld, A.rows, A.cols, eltype_code(A.dtype),
             1 if accumulate else 0, current_stream())
        return out

    def fill(self, dst, value):
        v = complex(value)
        call("fhip_fill", dst.ptr, eltype_code(dst.dtype), dst.ld, dst.rows, dst.cols, v.real, v.imag,
             current_stream())
        return dst

    def ghost_fill(self, dst, row0, col0, seed):
        """Entries (row0 .. row0+rows-1, col0 .. col0+cols-1) of a ghost matrix
        (0-based rows, 1-based columns) into dst."""
        call("fhip_ghost_randn", dst.ptr, eltype_code(dst.dtype), dst.ld, int(row0), dst.rows, int(col0),
             dst.cols, int(seed) & 0xFFFFFFFFFFFFFFFF, current_stream())
        return dst


# Columns per launch of the kernels that take one grid row per column.
_MAX_GRID_COLS = 65535


def _stream_of(queue):
    return None if queue is None else queue.stream


def _workspace(backend, work, need):
    """Pointer and size of the kernel workspace; a missing one is allocated
    here, which only the calls outside a sweep do."""
    if need == 0:
        return None, 0
    if work is None:
        work = backend.alloc_device(np.uint8, (need, 1))
        # Keep it alive until the stream has used it: the kernel is queued.
        _pending_workspaces.append(work)
        if len(_pending_workspaces) > 64:
            del _pending_workspaces[:32]
        return work.ptr, need
    if work.nbytes < need:
        raise ArgumentError(f"kernel workspace of {work.nbytes} bytes is smaller than the {need} needed")
    return work.ptr, work.nbytes


_pending_workspaces = []

_backend = None


def _array_backend():
    """The backend a DeviceArray's own updates launch through: the one already
    made, else the default device's."""
    return _backend if _backend is not None else hip_backend()


def hip_backend(device=0):
    """Funicular.hip_backend(): the HIP device backend on `device`."""
    global _backend
    if _backend is None or _backend.device != device:
        _backend = HIPBackend(device)
    return _backend
