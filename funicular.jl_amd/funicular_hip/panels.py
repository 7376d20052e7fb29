"""PanelMatrix and the movement of panels between tiers (src/panels.jl:1-555).

A PanelMatrix is an N x k matrix cut into column panels of width w (the last
one ragged). Each panel lives on one tier: the host tier (a block of a
page-locked slab), the ghost tier (regenerated on demand, here straight into
device memory), or -- under a plan built with device_resident=True -- the
device tier, where the whole matrix is one HBM allocation and sweeps run on
views of it without staging copies.

Panels and columns are numbered from 0 here; a ghost column's random stream is
keyed on its 1-based index in the full matrix, as in the reference.
"""
import numpy as np

from . import plan as _plan
from .device_api import ArgumentError, DimensionMismatch, current_stream
from .tiers import (DISK, GHOST, HOST, DeviceStorage, DeviceTier, DiskTier, GhostHome, GhostTier, HostStorage,
                    HostTier, storage)


class Panel:
    """src/panels.jl:1-13."""

    __slots__ = ("data", "tier", "home", "width", "dirty", "epoch", "pins", "hold")

    def __init__(self, data, tier, home, width):
        self.data = data
        self.tier = tier
        self.home = home
        self.width = width
        self.dirty = False
        self.epoch = 0
        self.pins = 0
        self.hold = None

    def __repr__(self):
        return f"Panel(width={self.width}, tier={self.tier!r}, pins={self.pins}, epoch={self.epoch})"


def panelstorage(panel):
    return storage(panel.data)


def isresident(panel):
    return panel.data is not None


def ondisk(panel):
    return isinstance(panel.tier, DiskTier)


def iscold(panel):
    return isinstance(panel.tier, (DiskTier, GhostTier))


def isreleased(panel):
    return panel.data is None and not iscold(panel)


def storage_eltype(panel):
    return panelstorage(panel).dtype


def host_storage_eltype(plan, T):
    """src/panels.jl:94-101."""
    T = np.dtype(T)
    if plan.host_eltype is None:
        return T
    S = plan.host_eltype
    if np.issubdtype(T, np.complexfloating) and not np.issubdtype(S, np.complexfloating):
        raise ArgumentError(
            f"a real host_eltype ({S}) would drop the imaginary parts of {T} panels; narrow a complex compute "
            "eltype to a complex storage eltype such as complex64")
    return S


def disk_storage_eltype(plan, T):
    """src/panels.jl:103-112."""
    S = host_storage_eltype(plan, T)
    if plan.disk_eltype is None:
        return S
    if plan.disk_eltype != S:
        raise ArgumentError(
            f"panels move between disk and host memory byte for byte, so disk_eltype ({plan.disk_eltype}) must "
            f"equal the host eltype ({S}); give host_eltype the same value, or drop disk_eltype")
    return S


class PanelMatrix:
    """Tall-skinny N x k matrix of compute eltype `dtype` in column panels of
    width w (src/panels.jl:22-92). Not an array: collect it with to_matrix()."""

    def __init__(self, N, k, *, plan, w=None, dtype=np.complex128, _panels=None, _device=None):
        T = np.dtype(dtype)
        if _panels is not None:
            self.N, self.k, self.w = int(N), int(k), int(w)
            self.panels = _panels
            self.plan = plan
            self.dtype = T
            self.store = None
            self.device = _device
            return
        if N <= 0:
            raise ArgumentError(f"N must be positive, got {N}")
        if k <= 0:
            raise ArgumentError(f"k must be positive, got {k}")
        if not plan.backend.supports_eltype(T):
            raise ArgumentError(f"PanelMatrix eltype must be one the backend computes in, got {T}")
        width = _plan.resolve_panel_width(plan, N, k, T, override=w)
        self.N, self.k, self.w = int(N), int(k), int(width)
        self.plan = plan
        self.dtype = T
        self.store = None
        self.device = None
        self.panels = []
        npan = _plan.cld(self.k, width)
        if plan.device_resident:
            nbytes = self.N * self.k * T.itemsize
            pool = plan.devicepool
            if pool.allocated + nbytes > pool.budget:
                raise ArgumentError(
                    f"a device-resident {self.N}×{self.k} {T} matrix needs {nbytes} bytes but the device budget "
                    f"has {pool.budget - pool.allocated} left")
            self.device = plan.backend.alloc_device(T, (self.N, self.k))
            pool.budget -= nbytes
            for j in range(npan):
                c0, c1 = j * width, min((j + 1) * width, self.k)
                p = Panel(DeviceStorage(self.device.view(None, (c0, c1))), DeviceTier(plan.backend), None, c1 - c0)
                self.panels.append(p)
            return
        S = host_storage_eltype(plan, T)
        D = disk_storage_eltype(plan, T)
        # With a scratch_dir each matrix is backed by its own file, the place
        # its panels go when host memory runs short (src/panels.jl:57-61).
        if plan.scratch_dir is not None:
            from .io import DiskHome, scratch_store
            self.store = scratch_store(plan, self.N, self.k, width, D, T)
        # Blocks are all sized for a full-width panel, the narrower last one
        # too, so any released block can take any panel (src/panels.jl:64-67).
        blockbytes = _plan.panel_bytes(self.N, width, S)
        try:
            for j in range(npan):
                c0, c1 = j * width, min((j + 1) * width, self.k)
                home = None if self.store is None else DiskHome(self.store, range(c0, c1))
                p = Panel(None, HOST, home, c1 - c0)
                self.panels.append(p)
                p.data = _plan.alloc_host_storage(plan, S, (self.N, c1 - c0), blockbytes)
                _plan.register_resident(plan, p)
        except BaseException:
            for p in self.panels:
                release(plan, p)
            if self.store is not None:
                from .io import close_store
                close_store(self.store)
            raise

    @classmethod
    def from_array(cls, A, *, plan, w=None):
        """PanelMatrix(A; plan, w): columns of A copied in panel by panel."""
        A = np.asarray(A)
        if A.ndim == 1:
            A = A.reshape(-1, 1)
        pm = cls(A.shape[0], A.shape[1], plan=plan, w=w, dtype=A.dtype)
        copyto(pm, A)
        return pm

    @property
    def shape(self):
        return (self.N, self.k)

    @property
    def resident(self):
        return self.device is not None

    def __repr__(self):
        s = f"PanelMatrix{{{self.dtype}}}({self.N}×{self.k}, w = {self.w}, {npanels(self)} panels"
        if all(isreleased(p) for p in self.panels) and not self.resident:
            s += ", released"
        else:
            if isghost(self):
                s += ", ghost"
            if self.resident:
                s += ", device resident"
            cold = sum(1 for p in self.panels if ondisk(p))
            if cold:
                s += f", {cold} on disk"
            S = host_storage_eltype(self.plan, self.dtype)
            if S != self.dtype and not self.resident:
                s += f", host storage {S}"
        return s + ")"


def npanels(pm):
    """The number of column panels of pm (src/panels.jl:131-137)."""
    return len(pm.panels)


def panelwidth(pm, j=None):
    """The panel width w of pm, or the width of panel j (the last one may be
    narrower; src/panels.jl:147-158)."""
    if j is None:
        return pm.w
    checkpanel(pm, j)
    return pm.panels[j].width


def panelrange(pm, j):
    """Columns of the full matrix that panel j holds, 0-based (src/panels.jl:161-170)."""
    checkpanel(pm, j)
    first = j * pm.w
    return range(first, first + pm.panels[j].width)


def checkpanel(pm, j):
    if not (0 <= j < npanels(pm)):
        raise IndexError(f"panel {j} is not among the {npanels(pm)} panels")


def panelindex(pm, col):
    return col // pm.w


def isghost(pm):
    return isinstance(pm.panels[0].home, GhostHome)


def assert_writable(pm):
    """src/ghosts.jl:81-84."""
    if isghost(pm):
        raise ArgumentError(
            "a ghost matrix cannot be written: its panels are recomputed from the seed whenever they are needed, "
            "so written values would not survive. Copy it into an ordinary PanelMatrix and write to that")


def assert_conformal(X, Y):
    if (X.N, X.k, X.w) != (Y.N, Y.k, Y.w):
        raise ArgumentError(
            f"the two panel matrices are partitioned differently: (N, k, w) = ({X.N}, {X.k}, {X.w}) against "
            f"({Y.N}, {Y.k}, {Y.w})")


def assert_same_panels(X, Y):
    if (X.k, X.w) != (Y.k, Y.w):
        raise ArgumentError(
            f"a panel-by-panel operation pairs panel j of one matrix with panel j of the other, which needs the same "
            f"(k, w); here they are ({X.k}, {X.w}) and ({Y.k}, {Y.w})")


def assert_same_rows(X, Y):
    if X.N != Y.N:
        raise ArgumentError(
            f"a row-block operation pairs equal row ranges of the two matrices, which needs equal N; here it is "
            f"{X.N} against {Y.N}")


def checkout_staging_buffers(plan, pm, dims, count):
    """src/panels.jl:114-122: buffers in the storage eltype for a narrowed host tier."""
    if pm.resident or (isghost(pm) and pm.panels[0].home.source.generator is None):
        return None
    S = host_storage_eltype(plan, pm.dtype)
    if S == pm.dtype:
        return None
    return _plan.checkout_device_buffers(plan, S, dims, count)


# --------------------------------------------------------------- movement --

def convert_staged(dst, src, plan, queue, event):
    """src/panels.jl:361-371."""
    backend = plan.backend
    if queue is not None:
        return backend.submit(queue, lambda: backend.convert_copy(dst, src))
    backend.convert_copy(dst, src)
    backend.sync_all()
    return event


def _cold_panel_error():
    return ArgumentError(
        "the panel to copy is on disk, not in host memory. prefetch it first, or copy it inside a sweep, whose "
        "producer stages each step's panels before the step's copies are issued")


def _not_resident_error():
    return ArgumentError(
        "the panel holds no data: free released it, or it was moved to a colder tier and not staged since")


def _ghost_fill(dst, panel, row0, queue, backend):
    home = panel.home
    col0 = home.cols[0]
    if queue is None:
        backend.ghost_fill(dst, row0, col0, home.source.seed)
        backend.sync_all()
        return None
    return backend.submit(queue, lambda: backend.ghost_fill(dst, row0, col0, home.source.seed))


def materialize(dst, panel, plan, queue=None, via=None):
    """Whole panel into a device buffer (src/panels.jl:373-382)."""
    backend = plan.backend
    if isinstance(panel.tier, GhostTier):
        if tuple(dst.shape) != (panel.home.source.N, panel.width):
            raise DimensionMismatch(f"device buffer is {dst.shape} but the ghost panel is "
                                    f"{(panel.home.source.N, panel.width)}")
        return _ghost_fill(dst, panel, 0, queue, backend)
    if isinstance(panel.tier, DiskTier):
        raise _cold_panel_error()
    if not isresident(panel):
        raise _not_resident_error()
    src = panelstorage(panel)
    if tuple(dst.shape) != tuple(src.shape):
        raise DimensionMismatch(f"the device buffer is {dst.shape} and the panel {src.shape}; copy a narrower "
                                "last panel into a column view of the buffer")
    if isinstance(panel.tier, DeviceTier):
        return backend.d2d(dst, src, queue=queue)
    if via is None:
        return backend.h2d(dst, src, queue=queue)
    event = backend.h2d(via, src, queue=queue)
    return convert_staged(dst, via, plan, queue, event)


def writeback(panel, src, plan, queue=None, via=None):
    """Device buffer into a whole panel, marking it dirty (src/panels.jl:384-398)."""
    backend = plan.backend
    if isinstance(panel.tier, GhostTier):
        raise ArgumentError("a ghost panel cannot be written")
    if isinstance(panel.tier, DiskTier):
        raise _cold_panel_error()
    if not isresident(panel):
        raise _not_resident_error()
    dst = panelstorage(panel)
    if tuple(dst.shape) != tuple(src.shape):
        raise DimensionMismatch(f"the device buffer is {src.shape} and the panel {dst.shape}; copy a narrower "
                                "last panel out of a column view of the buffer")
    if isinstance(panel.tier, DeviceTier):
        event = backend.d2d(dst, src, queue=queue)
    else:
        if via is not None:
            convert_staged(via, src, plan, queue, None)
            src = via
        event = backend.d2h(dst, src, queue=queue)
    panel.dirty = True
    panel.epoch += 1
    return event


def materialize_block(dst, panel, rows, plan, queue=None):
    """Rows of one panel into a device block (src/panels.jl:400-407)."""
    backend = plan.backend
    r0, r1 = rows
    if isinstance(panel.tier, GhostTier):
        if tuple(dst.shape) != (r1 - r0, panel.width):
            raise DimensionMismatch(f"device buffer block is {dst.shape} but rows {r0}:{r1} of the ghost panel are "
                                    f"{(r1 - r0, panel.width)}")
        return _ghost_fill(dst, panel, r0, queue, backend)
    if isinstance(panel.tier, DiskTier):
        raise _cold_panel_error()
    if not isresident(panel):
        raise _not_resident_error()
    st = panelstorage(panel)
    if isinstance(panel.tier, DeviceTier):
        src = st.view((r0, r1), None)
        return backend.d2d(dst, src, queue=queue)
    src = st[r0:r1, :]
    if tuple(dst.shape) != tuple(src.shape):
        raise DimensionMismatch(f"device buffer block is {dst.shape} but rows {r0}:{r1} of the host panel are {src.shape}")
    return backend.h2d(dst, src, queue=queue)


def writeback_block(panel, src, rows, plan, queue=None):
    """src/panels.jl:409-419."""
    backend = plan.backend
    r0, r1 = rows
    if isinstance(panel.tier, GhostTier):
        raise ArgumentError("a ghost panel cannot be written")
    if isinstance(panel.tier, DiskTier):
        raise _cold_panel_error()
    if not isresident(panel):
        raise _not_resident_error()
    st = panelstorage(panel)
    if isinstance(panel.tier, DeviceTier):
        event = backend.d2d(st.view((r0, r1), None), src, queue=queue)
    else:
        dst = st[r0:r1, :]
        if tuple(dst.shape) != tuple(src.shape):
            raise DimensionMismatch(f"device buffer block is {src.shape} but rows {r0}:{r1} of the host panel are "
                                    f"{dst.shape}")
        event = backend.d2h(dst, src, queue=queue)
    panel.dirty = True
    panel.epoch += 1
    return event


def materialize_rows(dst, pm, rows, queue=None, via=None):
    """A row block of the whole matrix, one 2-D copy per panel
    (src/panels.jl:432-445)."""
    landing = dst if via is None else via
    event = None
    for j in range(npanels(pm)):
        r = panelrange(pm, j)
        event = materialize_block(landing.view(None, (r.start, r.stop)), pm.panels[j], rows, pm.plan, queue=queue)
    if via is None:
        return event
    return convert_staged(dst, via, pm.plan, queue, event)


def writeback_rows(pm, src, rows, queue=None, via=None):
    """src/panels.jl:447-458."""
    leaving = src
    if via is not None:
        convert_staged(via, src, pm.plan, queue, None)
        leaving = via
    event = None
    for j in range(npanels(pm)):
        r = panelrange(pm, j)
        event = writeback_block(pm.panels[j], leaving.view(None, (r.start, r.stop)), rows, pm.plan, queue=queue)
    return event


# --------------------------------------------------------- tier changes --

def stage_host(panel, plan):
    """Bring a panel to where a sweep can copy from it and pin it there
    (src/panels.jl:475-505, src/ghosts.jl:105-122). A ghost panel on the
    device stream is generated by the copy itself, so staging it only pins
    it; one with a host generator is generated into a host block here, as
    the reference does, and from then on is a host panel until spilled."""
    if isinstance(panel.tier, (HostTier, DeviceTier)):
        if not isresident(panel):
            raise ArgumentError("this panel's storage was released (free, or a spill with nothing to bring it "
                                "back)")
        _plan.pin(plan, panel)
        return panel
    if isinstance(panel.tier, GhostTier):
        _plan.pin(plan, panel)
        source = panel.home.source
        if source.generator is None:
            return panel
        from .ghosts import generate_panel
        S = host_storage_eltype(plan, source.dtype)
        try:
            hs = _plan.alloc_host_storage(plan, S, (source.N, panel.width),
                                          _plan.panel_bytes(source.N, source.w, S))
        except BaseException:
            _plan.unpin(plan, panel)
            raise
        try:
            generate_panel(panel.home, hs.array)
        except BaseException:
            _plan.release_host_block(plan, hs.block)
            _plan.unpin(plan, panel)
            raise
        panel.data = hs
        panel.tier = HOST
        _plan.register_resident(plan, panel)
        return panel
    if isinstance(panel.tier, DiskTier):
        from .io import stage_host_disk
        return stage_host_disk(panel, plan)
    raise ArgumentError(f"unknown tier {panel.tier!r}")


def prefetch(pm, j):
    """src/panels.jl:460-473."""
    checkpanel(pm, j)
    panel = pm.panels[j]
    stage_host(panel, pm.plan)
    _plan.unpin(pm.plan, panel)
    return panel


def spill(plan, panel):
    """Give the host block back, writing the panel out first if it changed
    (src/panels.jl:507-520)."""
    hold = panel.hold
    if hold is not None:
        plan.backend.sync_event(hold)
    panel.hold = None
    home = panel.home
    if home is not None and not isinstance(home, GhostHome):
        from .io import cool
        cool(home, panel)
    release(plan, panel)
    panel.tier = GHOST if isinstance(home, GhostHome) else DISK


def evict(panel, plan):
    """src/panels.jl:531-536."""
    if iscold(panel):
        return None
    if isinstance(panel.tier, DeviceTier):
        raise ArgumentError("a device-resident panel has no colder tier to go to")
    if panel.home is None:
        raise ArgumentError(
            f"evicting a host panel needs a colder tier, and this plan has none (scratch_dir = "
            f"{plan.scratch_dir!r})")
    if panel.pins != 0:
        raise ArgumentError("the panel is pinned by a sweep in progress, which may still copy from its host "
                            "block; evict it after that sweep returns")
    spill(plan, panel)


def release(plan, panel):
    """src/panels.jl:538-543."""
    _plan.unregister_resident(plan, panel)
    if isinstance(panel.data, HostStorage):
        _plan.release_host_block(plan, panel.data.block)
    panel.data = None


def free(pm):
    """src/panels.jl:545-555."""
    for panel in pm.panels:
        release(pm.plan, panel)
        panel.tier = HOST
        panel.home = None
        panel.hold = None
        panel.pins = 0
    if pm.device is not None:
        pm.plan.devicepool.budget += pm.N * pm.k * pm.dtype.itemsize
        pm.device = None
    if pm.store is not None:
        from .io import close_store
        close_store(pm.store)


# ------------------------------------------------------------ host copies --

def copycols(dest, dcols, src, scols=None):
    """Whole columns into columns dcols of dest (src/panels.jl:211-304),
    from a dense host array or from columns scols of another PanelMatrix.
    Ranges are Python ranges (0-based)."""
    assert_writable(dest)
    dcols = _as_range(dcols)
    _assert_column_range(dcols, dest.k, "destination")
    if isinstance(src, PanelMatrix):
        if dest is src:
            raise ArgumentError("copycols needs distinct source and destination matrices")
        scols = _as_range(scols)
        _assert_column_range(scols, src.k, "source")
        if src.N != dest.N:
            raise DimensionMismatch(f"row counts differ: {src.N} in the source, {dest.N} in the destination")
        if len(dcols) != len(scols):
            raise DimensionMismatch(f"{len(scols)} source columns ({scols}) cannot fill {len(dcols)} destination "
                                    f"columns ({dcols})")
        dense = np.empty((src.N, len(scols)), dtype=src.dtype, order="F")
        _collect_columns(src, scols, dense)
        return copycols(dest, dcols, dense)
    A = np.asarray(src)
    if A.ndim == 1:
        A = A.reshape(-1, 1)
    if A.shape != (dest.N, len(dcols)):
        raise DimensionMismatch(
            f"the source is {A.shape[0]}×{A.shape[1]} but columns {dcols} of the destination are "
            f"{dest.N}×{len(dcols)}")
    if len(dcols) == 0:
        return dest
    backend = dest.plan.backend
    for j in range(npanels(dest)):
        pr = panelrange(dest, j)
        lo, hi = max(pr.start, dcols.start), min(pr.stop, dcols.stop)
        if lo >= hi:
            continue
        panel = dest.panels[j]
        sub = A[:, lo - dcols.start:hi - dcols.start]
        if dest.resident:
            block = np.asfortranarray(sub.astype(dest.dtype, copy=False))
            backend.h2d(panelstorage(panel).view(None, (lo - pr.start, hi - pr.start)), block)
        else:
            stage_host(panel, dest.plan)
            try:
                st = panelstorage(panel)
                st[:, lo - pr.start:hi - pr.start] = sub
            finally:
                _plan.unpin(dest.plan, panel)
        panel.dirty = True
        panel.epoch += 1
    return dest


def _as_range(r):
    if isinstance(r, range):
        if r.step != 1:
            raise ArgumentError("column ranges must be unit ranges")
        return r
    if isinstance(r, slice):
        return range(r.start or 0, r.stop)
    raise ArgumentError(f"a column range is a Python range, got {r!r}")


def _assert_column_range(cols, k, what):
    if not (cols.start >= 0 and cols.stop <= k and cols.start <= cols.stop):
        raise ArgumentError(f"columns {cols} are not among the {what}'s {k} columns")


def copyto(dest, src):
    """Whole-matrix copy (src/panels.jl:315-320)."""
    if isinstance(src, PanelMatrix):
        if (src.N, src.k) != (dest.N, dest.k):
            raise DimensionMismatch(f"copyto needs equal shapes; {src.N}×{src.k} into {dest.N}×{dest.k}")
        return copycols(dest, range(dest.k), src, range(src.k))
    return copycols(dest, range(dest.k), src)


class pooled_like:
    """A context manager giving a scratch PanelMatrix shaped like pm. Under a
    device-resident plan its storage is an N x k buffer checked out of the
    plan's device pool and checked back in on exit, so repeated calls reuse
    one allocation instead of allocating and freeing gigabytes each time;
    other plans get an ordinary PanelMatrix, freed on exit."""

    def __init__(self, pm):
        self.pm = pm
        self.buffer = None
        self.scratch = None

    def __enter__(self):
        pm = self.pm
        plan = pm.plan
        if not plan.device_resident:
            self.scratch = similar(pm)
            return self.scratch
        self.buffer = _plan.checkout_device_buffers(plan, pm.dtype, (pm.N, pm.k), 1)[0]
        panels = []
        for j in range(npanels(pm)):
            r = panelrange(pm, j)
            panels.append(Panel(DeviceStorage(self.buffer.view(None, (r.start, r.stop))), DeviceTier(plan.backend),
                                None, len(r)))
        self.scratch = PanelMatrix(pm.N, pm.k, plan=plan, w=pm.w, dtype=pm.dtype, _panels=panels,
                                   _device=self.buffer)
        return self.scratch

    def __exit__(self, *exc):
        if self.buffer is not None:
            for p in self.scratch.panels:
                p.data = None
            self.scratch.device = None
            _plan.checkin_device_buffers(self.pm.plan, [self.buffer])
        elif self.scratch is not None:
            free(self.scratch)
        self.scratch = None
        return False


def similar(pm, dtype=None, dims=None):
    """src/panels.jl:322-344: a new matrix on the same plan, keeping the panel
    width unless the new k is smaller."""
    T = pm.dtype if dtype is None else np.dtype(dtype)
    N, k = (pm.N, pm.k) if dims is None else (int(dims[0]), int(dims[1]))
    return PanelMatrix(N, k, plan=pm.plan, w=min(pm.w, k), dtype=T)


def _collect_columns(pm, cols, out):
    """Columns `cols` of pm into the host array out (N x len(cols))."""
    backend = pm.plan.backend
    for j in range(npanels(pm)):
        pr = panelrange(pm, j)
        lo, hi = max(pr.start, cols.start), min(pr.stop, cols.stop)
        if lo >= hi:
            continue
        panel = pm.panels[j]
        dst = out[:, lo - cols.start:hi - cols.start]
        # A ghost panel with a host generator is staged like a host panel:
        # generated into a host block, where it stays until evicted, as
        # Matrix(pm) leaves it in the reference (src/panels.jl:346-359).
        if isinstance(panel.tier, GhostTier) and panel.home.source.generator is None:
            tmp = backend.alloc_device(pm.dtype, (pm.N, hi - lo))
            # The fill and the blocking copy back share the calling thread's
            # stream, so the copy sees the fill without a wait of its own.
            backend.ghost_fill(tmp, 0, panel.home.cols[0] + (lo - pr.start), panel.home.source.seed)
            host = np.empty((pm.N, hi - lo), dtype=pm.dtype, order="F")
            backend.d2h(host, tmp)
            dst[...] = host
        elif pm.resident:
            host = np.empty((pm.N, hi - lo), dtype=pm.dtype, order="F")
            backend.d2h(host, panelstorage(panel).view(None, (lo - pr.start, hi - pr.start)))
            dst[...] = host
        else:
            stage_host(panel, pm.plan)
            try:
                dst[...] = panelstorage(panel)[:, lo - pr.start:hi - pr.start]
            finally:
                _plan.unpin(pm.plan, panel)


def to_matrix(pm, max_bytes=None):
    """Julia's Matrix(pm) (src/panels.jl:346-359)."""
    nbytes = pm.N * pm.k * pm.dtype.itemsize
    if max_bytes is not None and nbytes > max_bytes:
        raise ArgumentError(f"collecting a {pm.N}×{pm.k} {pm.dtype} PanelMatrix needs {nbytes} bytes but max_bytes "
                            f"is {int(max_bytes)}")
    A = np.empty((pm.N, pm.k), dtype=pm.dtype, order="F")
    _collect_columns(pm, range(pm.k), A)
    return A
