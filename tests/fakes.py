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
