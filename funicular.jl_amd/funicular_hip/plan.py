"""ResidencyPlan: the device buffer pool and the host block pool
(src/plan.jl:1-320), with the panel-width and row-block arithmetic that fixes
every sweep's schedule. The integer arithmetic here has to agree with the
reference to the bit, since it decides which columns and rows each step moves.
"""
import threading

import numpy as np

from .device_api import ArgumentError, SLAB_ALIGN, human_readable_bytes, slab_matrix
from .tiers import HostBlock, HostStorage

GiB = 1024 ** 3

# How far the sweep's producer runs ahead of the issue loop (src/plan.jl:7).
FEED_DEPTH = 2


def fld(a, b):
    return a // b


def cld(a, b):
    return -((-a) // b)


def align_up(n, a):
    return cld(n, a) * a


SCHEDULES = ("reference", "fused")


class DevicePool:
    """src/plan.jl:9-15: buffers by (eltype, dims), reused after checkin."""

    def __init__(self, budget):
        self.budget = int(budget)
        self.allocated = 0
        self.available = {}


class HostPool:
    """src/plan.jl:17-27: doubling slabs cut into 64-byte aligned blocks."""

    def __init__(self, budget):
        self.budget = int(budget)
        self.reserved = 0
        self.slabs = []
        self.cursors = []
        self.free = {}
        self.residents = []
        self.lock = threading.RLock()


class ResidencyPlan:
    """Owner of everything Funicular allocates (src/plan.jl:29-87).

    Keywords as in the reference: device_budget, host_budget, backend,
    scratch_dir, workspace_bytes, nbuffers, target_panel_bytes, panel_width,
    host_eltype, disk_eltype. Two keywords are this backend's additions:

    * `device_resident=True`: panels of matrices built from the plan live in
      HBM for their whole life, and sweeps over them skip the host tier (a
      DeviceTier).
    * `schedule`: the default step and transfer schedule of the panel BLAS.
      "reference" (the default) runs the reference's steps and copies one for
      one: cholqr2 as two (or three) passes of a Gram row sweep, the k x k
      Gram down to the host, a host Cholesky, the factor up and a solve row
      sweep; project as its p (p + 1) panel steps. "fused" lets the backend
      shorten them: the Cholesky on the device, the next pass's Gram formed
      inside the solve sweep off HBM, one solve launch over all rows of a
      resident matrix, and project by a resident G Q or by runs of held
      images. The results agree to rounding; the transfers do not.
    """

    def __init__(self, *, device_budget, host_budget, backend=None, scratch_dir=None, workspace_bytes=0,
                 nbuffers=2, target_panel_bytes=2 * GiB, panel_width=None, host_eltype=None,
                 disk_eltype=None, device_resident=False, schedule="reference"):
        if backend is None:
            from .device_api import hip_backend
            backend = hip_backend()
        dev = int(np.floor(device_budget))
        host = int(np.floor(host_budget))
        work = int(np.floor(workspace_bytes))
        if work < 0:
            raise ArgumentError(f"workspace_bytes must be nonnegative, got {work}")
        if nbuffers < 1:
            raise ArgumentError(f"a plan needs one staging buffer or more per operand; nbuffers = {nbuffers}")
        if target_panel_bytes < 1:
            raise ArgumentError(f"target_panel_bytes must be positive, got {target_panel_bytes}")
        if host <= 0:
            raise ArgumentError(f"host_budget must be positive, got {host}")
        if not dev > work:
            raise ArgumentError(
                f"the operator workspace ({human_readable_bytes(work)}) takes the whole device_budget "
                f"({human_readable_bytes(dev)}) and leaves no room for staging buffers; raise device_budget or "
                "lower workspace_bytes")
        width = None if panel_width is None else int(panel_width)
        if width is not None and width < 1:
            raise ArgumentError(f"panel_width must be positive, got {width}")
        for name, S in (("host_eltype", host_eltype), ("disk_eltype", disk_eltype)):
            if S is None:
                continue
            try:
                np.dtype(S)
            except TypeError:
                raise ArgumentError(f"{name} has to name a numeric eltype; {S!r} does not")
            if not backend.supports_eltype(S):
                raise ArgumentError(f"{name} {S} is not an eltype this backend stores")
        self.backend = backend
        self.device_budget = dev
        self.host_budget = host
        self.scratch_dir = None if scratch_dir is None else str(scratch_dir)
        self.workspace_bytes = work
        self.nbuffers = int(nbuffers)
        self.target_panel_bytes = int(np.floor(target_panel_bytes))
        self.panel_width = width
        self.host_eltype = None if host_eltype is None else np.dtype(host_eltype)
        self.disk_eltype = None if disk_eltype is None else np.dtype(disk_eltype)
        self.device_resident = bool(device_resident)
        if schedule not in SCHEDULES:
            raise ArgumentError(f"schedule must be one of {', '.join(map(repr, SCHEDULES))}, got {schedule!r}")
        self.schedule = schedule
        self.devicepool = DevicePool(dev - work)
        self.hostpool = HostPool(host)

    def __repr__(self):
        s = (f"ResidencyPlan({type(self.backend).__name__}, device {human_readable_bytes(self.device_budget)}, "
             f"host {human_readable_bytes(self.host_budget)}")
        if self.scratch_dir is not None:
            s += f", scratch {self.scratch_dir!r}"
        if self.device_resident:
            s += ", device resident"
        if self.schedule != "reference":
            s += f", {self.schedule} schedule"
        return s + ")"


def panel_bytes(N, w, T):
    """src/plan.jl:97."""
    return int(N) * int(w) * np.dtype(T).itemsize


def device_bytes_required(N, w, T, nbuffers=2, extra_buffers=0, workspace_bytes=0):
    """src/plan.jl:99-101."""
    return (int(nbuffers) + int(extra_buffers)) * panel_bytes(N, w, T) + int(workspace_bytes)


def choose_panel_width(N, k, T, *, device_budget, nbuffers=2, extra_buffers=0, workspace_bytes=0,
                       target_panel_bytes=2 * GiB):
    """The panel width: the widest for which nbuffers + extra_buffers panel
    buffers plus the workspace stay within the device budget, no wider than
    target_panel_bytes allows, and then balanced so the panels it gives are
    as equal as possible (src/plan.jl:103-123)."""
    if N <= 0:
        raise ArgumentError(f"N must be positive, got {N}")
    if k <= 0:
        raise ArgumentError(f"k must be positive, got {k}")
    buffers = int(nbuffers) + int(extra_buffers)
    if buffers < 1:
        raise ArgumentError(f"a width is chosen for one buffer or more; nbuffers + extra_buffers = {buffers}")
    column = int(N) * np.dtype(T).itemsize
    available = int(np.floor(device_budget)) - int(np.floor(workspace_bytes))
    per_buffer = fld(available, buffers)
    if per_buffer < column:
        raise ArgumentError(
            f"one column of {N} {np.dtype(T)} entries ({human_readable_bytes(column)}) does not fit a staging "
            f"buffer: the device_budget of {human_readable_bytes(int(np.floor(device_budget)))}, less "
            f"{human_readable_bytes(int(np.floor(workspace_bytes)))} of operator workspace, split over {buffers} "
            f"buffers gives {human_readable_bytes(per_buffer)} each. Raise device_budget, use fewer buffers, or a "
            "narrower compute eltype")
    w = min(int(k), max(1, fld(min(int(np.floor(target_panel_bytes)), per_buffer), column)))
    return cld(int(k), cld(int(k), w))


def row_block_height(N, k, w):
    """Height of a block of the row traversal, chosen so that a block of all k
    columns is about the size of one N x w panel (src/plan.jl:125-133)."""
    return min(max(fld(int(N) * int(w), int(k)), 1), int(N))


def resolve_panel_width(plan, N, k, T, override=None, extra_buffers=0):
    """src/plan.jl:135-157."""
    requested = plan.panel_width if override is None else override
    if requested is None:
        return choose_panel_width(N, k, T, device_budget=plan.device_budget, nbuffers=plan.nbuffers,
                                  extra_buffers=extra_buffers, workspace_bytes=plan.workspace_bytes,
                                  target_panel_bytes=plan.target_panel_bytes)
    w = int(requested)
    if not (1 <= w <= k):
        raise ArgumentError(f"a panel width for k = {k} lies between 1 and {k}; got {w}")
    needed = device_bytes_required(N, w, T, nbuffers=plan.nbuffers, extra_buffers=extra_buffers,
                                   workspace_bytes=plan.workspace_bytes)
    if needed > plan.device_budget:
        buffers = plan.nbuffers + int(extra_buffers)
        fits = choose_panel_width(N, k, T, device_budget=plan.device_budget, nbuffers=plan.nbuffers,
                                  extra_buffers=extra_buffers, workspace_bytes=plan.workspace_bytes,
                                  target_panel_bytes=plan.device_budget)
        raise ArgumentError(
            f"at panel width {w} the device would hold {buffers} buffers of "
            f"{human_readable_bytes(panel_bytes(N, w, T))} and {human_readable_bytes(plan.workspace_bytes)} of "
            f"operator workspace, {human_readable_bytes(needed)} in all, over the device_budget of "
            f"{human_readable_bytes(plan.device_budget)}: choose panel_width ≤ {fits}, or a larger budget")
    return w


def checkout_device_buffers(plan, T, dims, count):
    """src/plan.jl:159-175: `count` buffers of eltype T and shape dims, from the
    pool when it has them, otherwise allocated within the budget."""
    T = np.dtype(T)
    if not plan.backend.supports_eltype(T) and T not in (np.dtype(np.uint8), np.dtype(np.int32)):
        raise ArgumentError(f"backend {type(plan.backend).__name__} cannot compute in {T}")
    pool = plan.devicepool
    dims = (int(dims[0]), int(dims[1]))
    cache = pool.available.setdefault((T, dims), [])
    buffers = []
    nbytes = dims[0] * dims[1] * T.itemsize
    for _ in range(int(count)):
        if not cache:
            if pool.allocated + nbytes > pool.budget:
                raise ArgumentError(
                    f"no room in the device pool for a {dims[0]}×{dims[1]} {T} buffer "
                    f"({human_readable_bytes(nbytes)}): {human_readable_bytes(pool.allocated)} of its "
                    f"{human_readable_bytes(pool.budget)} is in use (the pool is the device_budget of "
                    f"{human_readable_bytes(plan.device_budget)} less the operator workspace). Use narrower "
                    "panels, fewer buffers or a larger device_budget")
            buffers.append(plan.backend.alloc_device(T, dims))
            pool.allocated += nbytes
        else:
            buffers.append(cache.pop())
    return buffers


def checkin_device_buffers(plan, buffers):
    """src/plan.jl:177-183."""
    for b in buffers:
        if b is None:
            continue
        plan.devicepool.available.setdefault((b.dtype, tuple(b.shape)), []).append(b)


def checkout_workspace(plan, nbytes):
    """A kernel workspace of at least nbytes, as a byte buffer out of the pool
    (None when nothing is needed)."""
    if nbytes <= 0:
        return None
    return checkout_device_buffers(plan, np.uint8, (align_up(int(nbytes), 256), 1), 1)[0]


def device_bytes_allocated(plan):
    """src/plan.jl:185."""
    return plan.devicepool.allocated


# ------------------------------------------------------------ host pool --

def acquire_host_block(plan, nbytes):
    """src/plan.jl:187-206: a block of nbytes (aligned), spilling unpinned
    panels that have a colder home until one fits."""
    from .panels import spill
    n = align_up(int(nbytes), SLAB_ALIGN)
    while True:
        block = take_host_block(plan, n)
        if block is not None:
            return block
        victim = claim_victim(plan, n)
        if victim is None:
            raise host_full_error(plan, n)
        try:
            spill(plan, victim)
        except BaseException:
            register_resident(plan, victim)
            raise


def take_host_block(plan, n):
    """src/plan.jl:208-225."""
    pool = plan.hostpool
    with pool.lock:
        reusable = pool.free.setdefault(n, [])
        if reusable:
            return reusable.pop()
        for i, slab in enumerate(pool.slabs):
            if pool.cursors[i] + n <= slab.nbytes:
                block = HostBlock(i, pool.cursors[i], n)
                pool.cursors[i] += n
                return block
        if not grow_host_pool(plan, n):
            return None
        i = len(pool.slabs) - 1
        block = HostBlock(i, pool.cursors[i], n)
        pool.cursors[i] += n
        return block


def grow_host_pool(plan, nbytes):
    """Slabs double so page-locking is amortized (src/plan.jl:227-238)."""
    pool = plan.hostpool
    remaining = pool.budget - pool.reserved
    if remaining < nbytes:
        return False
    slab = plan.backend.alloc_host_slab(min(remaining, max(int(nbytes), pool.reserved)))
    pool.slabs.append(slab)
    pool.cursors.append(0)
    pool.reserved += slab.nbytes
    return True


def claim_victim(plan, n):
    """The longest-resident unpinned panel holding a block of exactly n bytes
    (src/plan.jl:240-255)."""
    pool = plan.hostpool
    with pool.lock:
        for i, panel in enumerate(pool.residents):
            if panel.pins != 0:
                continue
            if not (isinstance(panel.data, HostStorage) and panel.data.block.nbytes == n):
                continue
            del pool.residents[i]
            return panel
        return None


def host_full_error(plan, nbytes):
    """src/plan.jl:257-266."""
    pool = plan.hostpool
    remaining = pool.budget - pool.reserved
    stranded = host_bytes_free(pool)
    detail = "" if stranded == 0 else (
        f"; another {human_readable_bytes(stranded)} is free but held in blocks of other sizes, which only "
        "blocks of those sizes can reuse")
    if plan.scratch_dir is None:
        return ArgumentError(
            f"out of host memory: a panel block of {human_readable_bytes(nbytes)} exceeds the "
            f"{human_readable_bytes(remaining)} still unreserved of the {human_readable_bytes(pool.budget)} "
            f"host_budget{detail}. Raise host_budget, make k or the panels smaller, or set a scratch_dir so that "
            "panels can move to disk")
    return ArgumentError(
        f"out of host memory, and every resident panel of this size is pinned, so no panel of that size is free "
        f"to spill: a block of {human_readable_bytes(nbytes)} exceeds the {human_readable_bytes(remaining)} "
        f"still unreserved of the {human_readable_bytes(pool.budget)} host_budget{detail}. A panel sweep pins the "
        f"panels of up to {FEED_DEPTH + 2} steps (one per matrix per step), and a row sweep (gram, cholqr2) pins "
        "every panel of its matrix; the host_budget has to cover that many")


def register_resident(plan, panel):
    """src/plan.jl:268-275."""
    if panel.home is None:
        return
    with plan.hostpool.lock:
        plan.hostpool.residents.append(panel)


def unregister_resident(plan, panel):
    """src/plan.jl:277-284."""
    pool = plan.hostpool
    with pool.lock:
        for i, p in enumerate(pool.residents):
            if p is panel:
                del pool.residents[i]
                break


def pin(plan, panel):
    """src/plan.jl:286-294."""
    with plan.hostpool.lock:
        panel.pins += 1


def unpin(plan, panel, hold=None):
    """src/plan.jl:296-302."""
    with plan.hostpool.lock:
        panel.pins -= 1
        if hold is not None:
            panel.hold = hold


def release_host_block(plan, block):
    """src/plan.jl:304-310."""
    with plan.hostpool.lock:
        plan.hostpool.free.setdefault(block.nbytes, []).append(block)


def host_bytes_reserved(pool):
    return pool.reserved


def host_bytes_free(pool):
    return sum(b.nbytes for blocks in pool.free.values() for b in blocks)


def host_bytes_in_use(pool):
    return sum(pool.cursors) - host_bytes_free(pool)


def alloc_host_storage(plan, T, dims, blockbytes=None):
    """src/plan.jl:316-320."""
    T = np.dtype(T)
    if blockbytes is None:
        blockbytes = int(dims[0]) * int(dims[1]) * T.itemsize
    block = acquire_host_block(plan, blockbytes)
    slab = plan.hostpool.slabs[block.slab]
    return HostStorage(slab_matrix(slab, T, dims, block.offset), slab, block)
