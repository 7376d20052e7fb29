"""A host-only stand-in for a device backend, so the plan's pool arithmetic
and the sweep's validation can be exercised on a machine without a GPU. It
computes nothing: its device arrays are bookkeeping records."""
import itertools

import numpy as np

from funicular_hip.device_api import DeviceArray, DeviceBackend, SLAB_ALIGN

_addresses = itertools.count(1 << 20, 1 << 12)


class _NumpySlab:
    def __init__(self, nbytes):
        self._buf = np.zeros(int(nbytes) + SLAB_ALIGN, dtype=np.uint8)
        base = self._buf.ctypes.data
        self.ptr = base + ((-base) % SLAB_ALIGN)
        self.nbytes = int(nbytes)


class HostOnlyBackend(DeviceBackend):
    def supports_eltype(self, T):
        return np.dtype(T) in (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.complex64),
                               np.dtype(np.complex128))

    def alloc_device(self, T, dims):
        T = np.dtype(T)
        rows, cols = int(dims[0]), int(dims[1])
        return DeviceArray(next(_addresses), (rows, cols), max(rows, 1), T, None)

    def alloc_host_slab(self, nbytes):
        return _NumpySlab(nbytes)


# ------------------------------------------------------- jittered backend --

import random
import threading
import time
import queue as _pyqueue

import scipy.linalg

_tls = threading.local()


class _Buffer:
    """Host memory standing in for one device allocation."""

    def __init__(self, nbytes):
        self.raw = np.zeros(int(nbytes) + 64, dtype=np.uint8)
        base = self.raw.ctypes.data
        self.offset = (-base) % 64
        self.base = base + self.offset


def _np(A):
    """The numpy view of a device array of the jittered backend."""
    if A.size == 0:
        return np.zeros(A.shape, dtype=A.dtype, order="F")
    buf = A.owner
    item = A.dtype.itemsize
    return np.ndarray(A.shape, dtype=A.dtype, buffer=buf.raw, offset=A.ptr - buf.base + buf.offset,
                      strides=(item, A.ld * item))


class _Worker:
    """One queue's thread: runs queued closures in order, each after a random
    sleep of up to `jitter` seconds."""

    def __init__(self, jitter, seed):
        self.jobs = _pyqueue.Queue()
        self.jitter = jitter
        self.rng = random.Random(seed)
        self.error = None
        self.thread = threading.Thread(target=self.run, daemon=True)
        self.thread.start()

    def run(self):
        while True:
            fn, done = self.jobs.get()
            if fn is None:
                done.set()
                return
            if self.jitter > 0:
                time.sleep(self.rng.random() * self.jitter)
            try:
                fn()
            except BaseException as err:  # surfaces at the next sync of this queue
                self.error = err
            done.set()


class JitterQueue:
    """A queue whose ordering comes only from its own FIFO and the events
    other queues wait on, as a HIP stream's; the backend's jitter varies the
    interleaving between queues from run to run (the reference's task-chain
    queues, src/device_api.jl:193-256)."""

    def __init__(self, backend):
        self.worker = _Worker(backend.jitter, backend.rng.random())
        self.tail = threading.Event()
        self.tail.set()
        backend._queues.add(self)

    def put(self, fn):
        done = threading.Event()
        self.worker.jobs.put((fn, done))
        self.tail = done
        return done

    def sync(self):
        self.tail.wait()
        err, self.worker.error = self.worker.error, None
        if err is not None:
            raise err

    def __del__(self):
        try:
            self.worker.jobs.put((None, threading.Event()))
        except Exception:
            pass


class JitteryBackend(DeviceBackend):
    """A device backend whose "device" is host memory and whose queues are
    worker threads that sleep a random time before each operation, so a
    missing wait between queues shows up as a wrong or varying result. The
    kernels are numpy and LAPACK, queued on the queue current in the calling
    thread (the one a sweep's submit makes current) or run at once outside a
    submit; the folded solve factor lives in its workspace, as the device's
    does, so a solve that runs before its factor is formed reads stale data.
    drop_waits=True ignores every wait_event: the planted ordering mistake
    the jitter is there to expose."""

    def __init__(self, jitter=2e-4, seed=0, drop_waits=False):
        import weakref
        self.jitter = float(jitter)
        self.rng = random.Random(seed)
        self.drop_waits = drop_waits
        self._queues = weakref.WeakSet()

    # -- memory --
    def supports_eltype(self, T):
        return np.dtype(T) in (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.complex64),
                               np.dtype(np.complex128), np.dtype(np.int32), np.dtype(np.uint8))

    def transfers_are_real(self):
        return False

    def alloc_device(self, T, dims):
        T = np.dtype(T)
        if len(dims) == 1:
            dims = (dims[0], 1)
        rows, cols = int(dims[0]), int(dims[1])
        buf = _Buffer(max(rows * cols * T.itemsize, 1))
        return DeviceArray(buf.base, (rows, cols), max(rows, 1), T, buf)

    def alloc_host_slab(self, nbytes):
        return _NumpySlab(nbytes)

    # -- queues --
    def make_queue(self):
        return JitterQueue(self)

    def record_event(self, queue):
        return queue.tail

    def wait_event(self, queue, event):
        if event is None or self.drop_waits:
            return None
        queue.put(event.wait)
        return None

    def submit(self, queue, work):
        previous = getattr(_tls, "queue", None)
        _tls.queue = queue
        try:
            work()
        finally:
            _tls.queue = previous
        return queue.tail

    def sync_queue(self, queue):
        queue.sync()

    def sync_event(self, event):
        if event is not None:
            event.wait()

    def sync_all(self):
        for q in list(self._queues):
            q.sync()

    def _launch(self, fn, queue=None):
        q = queue if queue is not None else getattr(_tls, "queue", None)
        if q is None:
            fn()
            return None
        return q.put(fn)

    def _copy(self, fn, queue=None):
        # A copy with a queue is asynchronous on it; without one it is
        # synchronous, on the current queue inside submit (after the work
        # already queued there), as the HIP backend's copies are.
        if queue is not None:
            return queue.put(fn)
        q = getattr(_tls, "queue", None)
        if q is None:
            fn()
        else:
            q.put(fn)
            q.sync()
        return None

    # -- copies --
    def h2d(self, dst, src, queue=None):
        src = np.asarray(src)
        if src.ndim == 1:
            src = src.reshape(-1, 1)
        if tuple(dst.shape) != tuple(src.shape):
            from funicular_hip.device_api import DimensionMismatch
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        return self._copy(lambda: np.copyto(_np(dst), src.astype(dst.dtype, copy=False)), queue)

    def d2h(self, dst, src, queue=None):
        if dst.ndim == 1:
            dst = dst.reshape(-1, 1)
        if tuple(dst.shape) != tuple(src.shape):
            from funicular_hip.device_api import DimensionMismatch
            raise DimensionMismatch(f"sizes differ: destination {dst.shape}, source {src.shape}")
        return self._copy(lambda: np.copyto(dst, _np(src).astype(dst.dtype, copy=False)), queue)

    def d2d(self, dst, src, queue=None):
        return self._copy(lambda: np.copyto(_np(dst), _np(src).astype(dst.dtype, copy=False)), queue)

    def to_host(self, A):
        out = np.empty(A.shape, dtype=A.dtype, order="F")
        self.d2h(out, A)
        return out

    def todevice(self, A):
        A = np.asfortranarray(A)
        if A.ndim == 1:
            A = A.reshape(-1, 1)
        dev = self.alloc_device(A.dtype, A.shape)
        np.copyto(_np(dev), A)
        return dev

    # -- kernels --
    def gram_workspace_bytes(self, m, ka, kb, T, hermitian=True, compute=None):
        return 64

    def trsm_workspace_bytes(self, k, T):
        return int(k) * int(k) * np.dtype(T).itemsize

    def dot_workspace_bytes(self, rows, cols):
        return 64

    def gram_accumulate(self, C, A, B, alpha, beta, work=None, compute=None):
        def run():
            a, b = _np(A), _np(B)
            if compute is not None:
                a, b = a.astype(compute), b.astype(compute)
            prod = (a.conj().T @ b).astype(C.dtype)
            c = _np(C)
            if complex(beta) == 0:
                c[...] = alpha * prod
            else:
                c[...] = alpha * prod + beta * c
        return self._launch(run)

    def potrf_upper(self, R, G, info):
        def run():
            g = np.triu(_np(G))
            g = g + np.triu(g, 1).conj().T
            potrf, = scipy.linalg.get_lapack_funcs(("potrf",), (g.astype(R.dtype),))
            c, bad = potrf(g.astype(R.dtype), lower=False, clean=True)
            _np(info)[0, 0] = bad
            if bad == 0:
                _np(R)[...] = c
        return self._launch(run)

    def _solve(self, Y, R, compute, adjoint):
        y = _np(Y)
        r = R.astype(compute) if compute is not None else R
        yy = y.astype(compute) if compute is not None else y
        if adjoint:
            x = scipy.linalg.solve_triangular(r.conj(), yy.T, trans="N", lower=False).T
        else:
            x = scipy.linalg.solve_triangular(r, yy.T, trans="T", lower=False).T
        y[...] = x.astype(Y.dtype)

    def _factor(self, work, k, T):
        return np.ndarray((k, k), dtype=T, buffer=work.owner.raw, offset=work.ptr - work.owner.base + work.owner.offset,
                          order="F")

    def rdiv_upper(self, Y, R, work=None, compute=None):
        T = np.dtype(R.dtype)

        def run():
            f = self._factor(work, R.rows, T)
            f[...] = _np(R)
            self._solve(Y, f.copy(), compute, False)
        return self._launch(run)

    def rdiv_upper_folded(self, Y, work, compute=None, adjoint=False):
        k = Y.cols
        T = np.dtype(Y.dtype)

        def run():
            self._solve(Y, self._factor(work, k, T).copy(), compute, adjoint)
        return self._launch(run)

    def rdiv_upper_adjoint(self, Y, R, work=None):
        T = np.dtype(R.dtype)

        def run():
            f = self._factor(work, R.rows, T)
            f[...] = _np(R)
            self._solve(Y, f.copy(), None, True)
        return self._launch(run)

    def rightmul_gemm(self, dst, A, C):
        return self._launch(lambda: np.copyto(_np(dst), _np(A) @ _np(C)))

    def rightmul_gemm_scaled(self, dst, A, C, alpha, beta):
        def run():
            d = _np(dst)
            prod = alpha * (_np(A) @ _np(C))
            d[...] = prod if complex(beta) == 0 else prod + beta * d
        return self._launch(run)

    def convert_copy(self, dst, src, stream=None):
        return self._launch(lambda: np.copyto(_np(dst), _np(src).astype(dst.dtype)))

    def axpby(self, Y, X, a, b):
        return self._launch(lambda: np.copyto(_np(Y), a * _np(X) + b * _np(Y)))

    def scale(self, Y, alpha):
        return self._launch(lambda: np.copyto(_np(Y), _np(Y) * alpha))

    def affine(self, Y, alpha, beta):
        return self._launch(lambda: np.copyto(_np(Y), _np(Y) * alpha + beta))

    def fill(self, dst, value):
        return self._launch(lambda: _np(dst).fill(value))

    def dot_accumulate(self, out, X, Y=None, accumulate=False, work=None):
        def run():
            x = _np(X)
            v = np.vdot(x, x) if Y is None else np.vdot(x, _np(Y))
            o = _np(out)
            o[0, 0] = (o[0, 0] if accumulate else 0) + v
        return self._launch(run)

    def rowdot(self, out, A, B, accumulate=False):
        def run():
            v = np.real(np.sum(_np(A).conj() * _np(B), axis=1))
            o = _np(out)[:A.rows, 0]
            o[...] = (o if accumulate else 0) + v
        return self._launch(run)
