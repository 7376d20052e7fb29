"""GhostPanels: a matrix that is never stored; each panel is produced from
its column indices and a seed whenever a sweep needs it (src/ghosts.jl:1-127).

Two kinds of generator:

* A host generator, as the reference has it (GhostPanels(f!, T, N, k),
  src/ghosts.jl:53-63), and the default: generator(dst, rng, col) fills the
  host column dst of 1-based index col, with rng a numpy Generator keyed by
  (seed, col); the producer thread runs it into a host block in the plan's
  host eltype, and the sweep uploads that block (src/ghosts.jl:97-122).
  randn_column is the reference's default generator in this form, and
  GhostPanels(T, N, k) uses it.
* The device stream, named by passing device_randn as the generator:
  fhip_ghost_randn writes the panel straight into the device buffer the sweep
  computes on, so a ghost matrix costs neither host memory nor host-to-device
  bandwidth. It is this library's own counter-based stream (SplitMix64 keys,
  a 256-layer ziggurat).

Either way column col's values depend on (seed, col) only, so any panel
width and any order give the same matrix. Neither reproduces the numbers of
Julia's randn! on Xoshiro(hash(col, hash(seed))): that stream needs Julia's
own hash, seeding and ziggurat, which are absent here.
"""
import numpy as np

from . import plan as _plan
from .device_api import ArgumentError
from .panels import Panel, PanelMatrix
from .tiers import GHOST, GhostHome, GhostSource


def randn_column(dst, rng, col):
    """Standard normal entries into the host column dst; a complex eltype gets
    real and imaginary parts of variance 1/2 each (src/ghosts.jl:69-77)."""
    n = dst.shape[0]
    if np.iscomplexobj(dst):
        dst[...] = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)
    else:
        dst[...] = rng.standard_normal(n)


class DeviceStream:
    """The generator that selects the device stream (fhip_ghost_randn); its
    one instance is device_randn."""

    __slots__ = ()

    def __repr__(self):
        return "device_randn"


device_randn = DeviceStream()


def column_rng(seed, col):
    """The numpy Generator a host generator gets for 1-based column col."""
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([int(seed) & 0xFFFFFFFFFFFFFFFF, int(col)])))


def generate_panel(home, dst):
    """The columns of one ghost panel into the host array dst, one generator
    call per column (src/ghosts.jl:97-103)."""
    source = home.source
    first, last = home.cols
    for c, col in enumerate(range(first, last + 1)):
        source.generator(dst[:, c], column_rng(source.seed, col), col)
    return dst


def GhostPanels(*args, plan, seed=0, w=None, generator=None):
    """An N x k ghost matrix of compute eltype dtype: GhostPanels(f, dtype,
    N, k) (or generator=f) with a host generator f(dst, rng, col), or with
    f = device_randn on the device stream; GhostPanels(dtype, N, k) takes
    randn_column, as the reference's GhostPanels(T, N, k) takes randn!
    (src/ghosts.jl:53-63)."""
    if len(args) == 4:
        if generator is not None:
            raise ArgumentError("give the column generator either first or as generator=, not both")
        generator, dtype, N, k = args
    elif len(args) == 3:
        dtype, N, k = args
    else:
        raise ArgumentError("GhostPanels takes (dtype, N, k) or (generator, dtype, N, k)")
    if generator is None:
        generator = randn_column
    if generator is not device_randn and not callable(generator):
        raise ArgumentError(f"a column generator is called as generator(dst, rng, col); got {generator!r}")
    if N <= 0:
        raise ArgumentError(f"a ghost matrix needs at least one row; N = {N}")
    if k <= 0:
        raise ArgumentError(f"a ghost matrix needs at least one column; k = {k}")
    try:
        T = np.dtype(dtype)
    except TypeError:
        raise ArgumentError(f"{dtype!r} is not a concrete eltype")
    if not plan.backend.supports_eltype(T):
        raise ArgumentError(f"the plan's backend does not compute in {T}")
    width = _plan.resolve_panel_width(plan, N, k, T, override=w)
    host = None if generator is device_randn else generator
    source = GhostSource(int(seed) & 0xFFFFFFFFFFFFFFFF, int(N), int(k), width, host, T)
    panels = []
    for j in range(_plan.cld(int(k), width)):
        c0, c1 = j * width, min((j + 1) * width, int(k))
        # 1-based column indices in the full matrix, as the stream is keyed.
        panels.append(Panel(None, GHOST, GhostHome(source, (c0 + 1, c1)), c1 - c0))
    return PanelMatrix(N, k, plan=plan, w=width, dtype=T, _panels=panels)
