"""The disk tier (src/io.jl, src/tiers.jl:41-72, src/panels.jl:487-536,
ext/FunicularHDF5Ext.jl:21-75).

One file per PanelMatrix holds the whole N x k matrix. The reference keeps it
as an HDF5 dataset "panels" of shape (N, k), chunked (N, w), with attributes
funicular_format = 1, panel_width, storage_eltype and compute_eltype. HDF5 is
not available to this backend, so the file keeps the same layout and the same
attributes in a raw form:

    bytes [0, 4096)     b"FUNICULAR-PANELS" then a JSON object with the
                        attributes above plus N, k and "dataset": "panels",
                        zero padded
    bytes [4096, ...)   the matrix, column major, in the storage eltype

A panel's columns are contiguous in column-major order, so one panel is one
contiguous region at 4096 + N * first_column * itemsize, and the chunk is the
unit of I/O exactly as in the HDF5 layout. A panel's region moves as parallel
64 MiB preads or pwrites straight into (out of) the panel's page-locked host
block, on the sweep's producer thread, and DISK_LOCK serializes the panels as
the reference serializes its HDF5 calls.
"""
import itertools
import json
import os
import threading

import numpy as np

from . import plan as _plan
from .device_api import ArgumentError, DimensionMismatch
from .tiers import DISK, HOST, DiskTier

MAGIC = b"FUNICULAR-PANELS"
FORMAT = 1
HEADER_BYTES = 4096
DATASET = "panels"

# File access is serialized by this lock. It is held around a single panel's
# read or write and released before the producer blocks on a free slot
# (src/io.jl:63-66).
DISK_LOCK = threading.RLock()
SCRATCH_COUNTER = itertools.count()

# What the cold tier can hold, by the reference's eltype names (src/io.jl:70-73).
STORAGE_ELTYPES = {"Float32": np.dtype(np.float32), "Float64": np.dtype(np.float64),
                   "ComplexF32": np.dtype(np.complex64), "ComplexF64": np.dtype(np.complex128)}
_ELTYPE_NAMES = {v: k for k, v in STORAGE_ELTYPES.items()}


def eltype_name(T):
    T = np.dtype(T)
    if T not in _ELTYPE_NAMES:
        raise ArgumentError(f"{T} has no place on disk; panel files hold {', '.join(sorted(STORAGE_ELTYPES))}. "
                            "Set the plan's storage eltypes to one of them, or keep the compute eltype")
    return _ELTYPE_NAMES[T]


def storage_eltype_named(name):
    """src/io.jl:75-78."""
    if name in STORAGE_ELTYPES:
        return STORAGE_ELTYPES[name]
    raise ArgumentError(f"the panel file names its eltype {name!r}; readable ones: "
                        f"{', '.join(sorted(STORAGE_ELTYPES))}")


# ------------------------------------------------- the file format itself --

class RawStore:
    """An open panel file: its descriptor and the layout its header records."""

    __slots__ = ("fd", "path", "N", "k", "w", "stored", "computed", "writable")

    def __init__(self, fd, path, N, k, w, stored, computed, writable):
        self.fd = fd
        self.path = path
        self.N, self.k, self.w = N, k, w
        self.stored, self.computed = stored, computed
        self.writable = writable


def open_panel_store(path, mode, *, N=0, k=0, w=0, stored=np.complex128, computed=np.complex128):
    """The RawStore over the panel file at path (src/io.jl:5-18,
    ext/FunicularHDF5Ext.jl:21-42). "w" truncates or makes the file and
    sizes it for N x k entries of `stored`, chunked w columns at a time; "r"
    and "r+" take N, k, w and the eltypes from the file's header."""
    if mode not in ("w", "r", "r+"):
        raise ArgumentError(f'mode {mode!r}: panel files open as "w", "r" or "r+"')
    if mode == "w":
        N, k, w = int(N), int(k), int(w)
        if N <= 0 or k <= 0 or w <= 0:
            raise ArgumentError(f"a panel store needs positive N, k and w, got {N}, {k}, {w}")
        stored, computed = np.dtype(stored), np.dtype(computed)
        header = {"funicular_format": FORMAT, "dataset": DATASET, "N": N, "k": k, "panel_width": min(w, k),
                  "storage_eltype": eltype_name(stored), "compute_eltype": eltype_name(computed),
                  "order": "column-major"}
        blob = MAGIC + json.dumps(header).encode()
        if len(blob) > HEADER_BYTES:
            raise ArgumentError("panel store header does not fit its 4096 bytes")
        fd = os.open(path, os.O_RDWR | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            _pwrite_all(fd, memoryview(blob.ljust(HEADER_BYTES, b"\0")), 0)
            os.ftruncate(fd, HEADER_BYTES + N * k * stored.itemsize)
        except BaseException:
            os.close(fd)
            raise
        return RawStore(fd, str(path), N, k, min(w, k), stored, computed, True)
    fd = os.open(path, os.O_RDONLY if mode == "r" else os.O_RDWR)
    try:
        raw = os.pread(fd, HEADER_BYTES, 0)
        if not raw.startswith(MAGIC):
            raise ArgumentError(f"{path!r} does not start with the panel-file magic: not one Funicular wrote")
        header = json.loads(raw[len(MAGIC):].rstrip(b"\0").decode())
        if header.get("funicular_format") != FORMAT:
            raise ArgumentError(f"the file at {path!r} is panel format {header.get('funicular_format')}, and this "
                                f"reader knows format {FORMAT}")
        N, k = int(header["N"]), int(header["k"])
        stored = storage_eltype_named(header["storage_eltype"])
        computed = storage_eltype_named(header.get("compute_eltype", header["storage_eltype"]))
        w = min(max(int(header.get("panel_width", k)), 1), k)
        size = os.fstat(fd).st_size
        if size < HEADER_BYTES + N * k * stored.itemsize:
            raise ArgumentError(f"the file at {path!r} is {size} bytes, short of the {N}×{k} {stored} matrix its "
                                "header describes")
    except BaseException:
        os.close(fd)
        raise
    return RawStore(fd, str(path), N, k, w, stored, computed, mode == "r+")


def close_panel_store(handle):
    os.close(handle.fd)


def panel_store_layout(handle):
    """(N, k, w, stored, computed) as the file records them (src/io.jl:27-32)."""
    return handle.N, handle.k, handle.w, handle.stored, handle.computed


def _region(handle, cols, arr):
    cols = range(cols.start, cols.stop)
    if not (0 <= cols.start <= cols.stop <= handle.k):
        raise ArgumentError(f"columns {cols} are outside the store's {handle.k} columns")
    if arr.dtype != handle.stored or arr.shape != (handle.N, len(cols)):
        raise DimensionMismatch(f"a {arr.shape} {arr.dtype} host array cannot hold columns {cols} of a "
                                f"{handle.N}×{handle.k} {handle.stored} store")
    if not arr.flags.f_contiguous:
        raise ArgumentError("panel I/O needs a column-major contiguous host array")
    offset = HEADER_BYTES + handle.N * cols.start * handle.stored.itemsize
    return memoryview(arr.reshape(-1, order="F").view(np.uint8)), offset


def _pwrite_all(fd, mv, offset):
    done = 0
    while done < len(mv):
        n = os.pwrite(fd, mv[done:], offset + done)
        if n <= 0:
            raise OSError("short write to a panel store")
        done += n


def _pread_all(fd, mv, offset):
    done = 0
    while done < len(mv):
        n = os.preadv(fd, [mv[done:]], offset + done)
        if n <= 0:
            raise OSError("a panel store ended before the region read from it")
        done += n


# A panel is gigabytes at the configurations' sizes, and one thread copying
# it through the page cache runs at a few GB/s. The region is cut into pieces
# that a small pool moves in parallel (the system calls release the GIL).
IO_PIECE_BYTES = 64 << 20
IO_THREADS = 8
_io_pool = None


def _parallel(fn, fd, mv, offset):
    global _io_pool
    if len(mv) <= IO_PIECE_BYTES:
        fn(fd, mv, offset)
        return
    if _io_pool is None:
        import concurrent.futures
        _io_pool = concurrent.futures.ThreadPoolExecutor(max_workers=IO_THREADS, thread_name_prefix="panel-io")
    pieces = [_io_pool.submit(fn, fd, mv[a:a + IO_PIECE_BYTES], offset + a)
              for a in range(0, len(mv), IO_PIECE_BYTES)]
    for f in pieces:
        f.result()


def read_panel_region(handle, dst, cols):
    """Reads columns `cols` of the stored matrix into the host array dst, of
    the store's eltype and exactly their shape (src/io.jl:34-40)."""
    mv, offset = _region(handle, cols, dst)
    _parallel(_pread_all, handle.fd, mv, offset)
    return dst


def write_panel_region(handle, src, cols):
    """Writes src into columns `cols` of the stored matrix (src/io.jl:42-47)."""
    if not handle.writable:
        raise ArgumentError(f"the store at {handle.path!r} is open for reading only")
    mv, offset = _region(handle, cols, np.asfortranarray(src))
    _parallel(_pwrite_all, handle.fd, mv, offset)
    return src


# ----------------------------------------------------------- the tier ----

class PanelStore:
    """One file behind one PanelMatrix (src/tiers.jl:41-62)."""

    __slots__ = ("path", "handle", "N", "k", "w", "stored", "computed", "readonly", "scratch", "reads", "writes")

    def __init__(self, path, handle, N, k, w, stored, computed, readonly, scratch):
        self.path = path
        self.handle = handle
        self.N, self.k, self.w = N, k, w
        self.stored, self.computed = stored, computed
        self.readonly = readonly
        self.scratch = scratch
        self.reads = 0
        self.writes = 0

    def __repr__(self):
        s = f"PanelStore({self.path!r}, {self.N}×{self.k} {eltype_name(self.stored)}, w = {self.w}"
        if self.readonly:
            s += ", read only"
        if self.scratch:
            s += ", scratch"
        return s + ")"


class DiskHome:
    """Where a cold panel lives: the matrix's store and the panel's columns in
    it (src/tiers.jl:66-71)."""

    __slots__ = ("store", "cols")

    def __init__(self, store, cols):
        self.store = store
        self.cols = cols


def open_store(path, mode, *, N=0, k=0, w=0, stored=np.complex128, computed=np.complex128, scratch=False):
    """src/io.jl:80-90."""
    if mode == "w":
        eltype_name(stored)
    with DISK_LOCK:
        handle = open_panel_store(path, mode, N=N, k=k, w=w, stored=stored, computed=computed)
        layout = panel_store_layout(handle)
    return PanelStore(str(path), handle, *layout, mode == "r", scratch)


def scratch_store(plan, N, k, w, S, T):
    """A scratch file of the plan's own, removed when the store closes
    (src/io.jl:92-97)."""
    os.makedirs(plan.scratch_dir, exist_ok=True)
    serial = next(SCRATCH_COUNTER)
    path = os.path.join(plan.scratch_dir, f"funicular-{os.getpid()}-{serial}.panels")
    return open_store(path, "w", N=N, k=k, w=w, stored=S, computed=T, scratch=True)


def close_store(store):
    """src/io.jl:99-109."""
    if store.handle is None:
        return None
    handle = store.handle
    store.handle = None
    with DISK_LOCK:
        close_panel_store(handle)
    if store.scratch:
        try:
            os.remove(store.path)
        except FileNotFoundError:
            pass
    return None


def read_panel(home, dst):
    """src/io.jl:111-120."""
    store = home.store
    if store.handle is None:
        raise ArgumentError(f"reading a panel from {store.path!r}, which free or save has closed already")
    with DISK_LOCK:
        read_panel_region(store.handle, dst, home.cols)
        store.reads += 1
    return dst


def write_panel(home, src):
    """src/io.jl:122-132."""
    store = home.store
    if store.handle is None:
        raise ArgumentError(f"writing a panel back to {store.path!r}, which free or save has closed already")
    if store.readonly:
        raise ArgumentError(f"a modified panel cannot go back to {store.path!r}: the matrix was loaded read only. "
                            "Load it with readonly=False to let operations change the file")
    with DISK_LOCK:
        write_panel_region(store.handle, src, home.cols)
        store.writes += 1
    return src


def disk_reads(pm):
    """Panel reads from pm's disk store so far, 0 without a store
    (src/io.jl:134-144)."""
    return 0 if pm.store is None else pm.store.reads


def disk_writes(pm):
    """Panel writes to pm's disk store so far, 0 without a store
    (src/io.jl:134-144)."""
    return 0 if pm.store is None else pm.store.writes


def stage_host_disk(panel, plan):
    """Reads a disk panel into a fresh page-locked host block and pins it
    there (src/panels.jl:487-505). On a sweep this runs on the producer
    thread, one step ahead of the pipeline."""
    home = panel.home
    _plan.pin(plan, panel)
    store = home.store
    dims = (store.N, len(home.cols))
    try:
        st = _plan.alloc_host_storage(plan, store.stored, dims, _plan.panel_bytes(store.N, store.w, store.stored))
    except BaseException:
        _plan.unpin(plan, panel)
        raise
    try:
        read_panel(home, st.array)
    except BaseException:
        _plan.release_host_block(plan, st.block)
        _plan.unpin(plan, panel)
        raise
    panel.data = st
    panel.tier = HOST
    _plan.register_resident(plan, panel)
    return panel


def cool(home, panel):
    """Writes a changed panel back before its host block goes (src/panels.jl:522-527)."""
    if not panel.dirty:
        return None
    from .panels import panelstorage
    write_panel(home, panelstorage(panel))
    panel.dirty = False
    return None


def save(pm, path):
    """A panel file at `path` holding pm's entries, converted to the plan's
    disk storage eltype; pm itself does not move (src/io.jl:146-172)."""
    from .panels import _collect_columns, disk_storage_eltype, npanels, panelrange
    S = disk_storage_eltype(pm.plan, pm.dtype)
    store = open_store(path, "w", N=pm.N, k=pm.k, w=pm.w, stored=S, computed=pm.dtype)
    try:
        for j in range(npanels(pm)):
            cols = panelrange(pm, j)
            host = np.empty((pm.N, len(cols)), dtype=pm.dtype, order="F")
            _collect_columns(pm, cols, host)
            write_panel(DiskHome(store, cols), np.asfortranarray(host.astype(S, copy=False)))
    finally:
        close_store(store)
    return path


def load(path, *, plan, readonly=False, dtype=None):
    """A PanelMatrix backed by the panel file at `path` (src/io.jl:174-212).
    Its panels live in that file: operations that change a panel change the
    file, and with readonly=True such a change raises instead."""
    from .panels import Panel, PanelMatrix, host_storage_eltype
    if not os.path.isfile(path):
        raise ArgumentError(f"load: no file at {path!r}")
    store = open_store(path, "r" if readonly else "r+")
    try:
        T = store.computed if dtype is None else np.dtype(dtype)
        S = host_storage_eltype(plan, T)
        if S != store.stored:
            raise ArgumentError(f"{path!r} stores {eltype_name(store.stored)} entries and the plan keeps {S} off the "
                                f"device; load it with host_eltype = {store.stored}")
        if plan.device_resident:
            raise ArgumentError("a device-resident plan keeps its matrices in HBM; load the file with a plan "
                                "that streams (device_resident=False)")
        # The width is the file's; the call below only checks that the plan
        # can stage panels of it.
        _plan.resolve_panel_width(plan, store.N, store.k, T, override=store.w)
        panels = []
        for j in range(_plan.cld(store.k, store.w)):
            cols = range(j * store.w, min((j + 1) * store.w, store.k))
            panels.append(Panel(None, DISK, DiskHome(store, cols), len(cols)))
        pm = PanelMatrix(store.N, store.k, plan=plan, w=store.w, dtype=T, _panels=panels)
        pm.store = store
        return pm
    except BaseException:
        close_store(store)
        raise


__all__ = ["DiskTier", "PanelStore", "DiskHome", "open_store", "close_store", "save", "load", "disk_reads",
           "disk_writes"]
