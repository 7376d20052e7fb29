"""Tier types and tier handles (src/tiers.jl:1-92)."""


class Tier:
    pass


class HostTier(Tier):
    def __repr__(self):
        return "host"


class DiskTier(Tier):
    def __repr__(self):
        return "disk"


class GhostTier(Tier):
    """A panel that is computed from its column indices whenever it is needed
    instead of being kept (src/tiers.jl:8-11)."""

    def __repr__(self):
        return "ghost"


class DeviceTier(Tier):
    """A panel that lives in device memory for its whole life. The reference
    declares this tier (src/tiers.jl:13-17) without using it; here it is the
    resident mode of a plan built with device_resident=True."""

    def __init__(self, backend=None):
        self.backend = backend

    def __repr__(self):
        return f"device({type(self.backend).__name__})"


HOST = HostTier()
DISK = DiskTier()
GHOST = GhostTier()


class HostBlock:
    """src/tiers.jl:23-27."""

    __slots__ = ("slab", "offset", "nbytes")

    def __init__(self, slab, offset, nbytes):
        self.slab = slab
        self.offset = offset
        self.nbytes = nbytes

    def __eq__(self, other):
        return isinstance(other, HostBlock) and (self.slab, self.offset, self.nbytes) == (
            other.slab, other.offset, other.nbytes)

    def __hash__(self):
        return hash((self.slab, self.offset, self.nbytes))

    def __repr__(self):
        return f"HostBlock(slab={self.slab}, offset={self.offset}, nbytes={self.nbytes})"


class HostStorage:
    """A host-resident panel: its array view, the slab that roots it and the
    block to give back (src/tiers.jl:29-36)."""

    __slots__ = ("array", "slab", "block")

    def __init__(self, array, slab, block):
        self.array = array
        self.slab = slab
        self.block = block


class DeviceStorage:
    """A device-resident panel: a view of the matrix's one HBM allocation."""

    __slots__ = ("array",)

    def __init__(self, array):
        self.array = array


def storage(handle):
    """src/tiers.jl:38-39."""
    if isinstance(handle, (HostStorage, DeviceStorage)):
        return handle.array
    return handle


class GhostSource:
    """Generator, seed and shape shared by every panel of a ghost matrix
    (src/tiers.jl:76-82). generator None is the device fill; otherwise it is
    a host function called per column as generator(dst, rng, col)."""

    __slots__ = ("seed", "N", "k", "w", "generator", "dtype")

    def __init__(self, seed, N, k, w, generator=None, dtype=None):
        self.seed = seed
        self.N = N
        self.k = k
        self.w = w
        self.generator = generator
        self.dtype = dtype

    def __repr__(self):
        return f"GhostSource(seed = {self.seed:#x}, {self.N}×{self.k}, w = {self.w})"


class GhostHome:
    """src/tiers.jl:84-87: the source and the 1-based columns of one panel."""

    __slots__ = ("source", "cols")

    def __init__(self, source, cols):
        self.source = source
        self.cols = cols  # (first, last) 1-based inclusive
