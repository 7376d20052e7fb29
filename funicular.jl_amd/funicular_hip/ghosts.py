"""GhostPanels: a matrix whose panels are generated on demand and never stored
(src/ghosts.jl:1-127).

The reference generates a ghost column on the host producer task with randn!
on a Xoshiro seeded from (col, seed) and then uploads it. Here the column is
generated by the device fill (fhip_ghost_randn) straight into whatever device
buffer the sweep copies it to, so a ghost matrix costs neither host memory nor
host-to-device bandwidth. Its contract is the reference's: column col's values
depend on (seed, col) only, so any panel width and any order regenerate the
same matrix, and entries are standard normal (complex parts of variance 1/2).
"""
import numpy as np

from . import plan as _plan
from .device_api import ArgumentError
from .panels import Panel, PanelMatrix
from .tiers import GHOST, GhostHome, GhostSource


def GhostPanels(dtype, N, k, *, plan, seed=0, w=None):
    """An N x k ghost matrix of compute eltype dtype (src/ghosts.jl:53-63)."""
    if N <= 0:
        raise ArgumentError(f"N must be positive, got {N}")
    if k <= 0:
        raise ArgumentError(f"k must be positive, got {k}")
    try:
        T = np.dtype(dtype)
    except TypeError:
        raise ArgumentError(f"GhostPanels eltype must be concrete, got {dtype}")
    if not plan.backend.supports_eltype(T):
        raise ArgumentError(f"GhostPanels eltype must be one the backend computes in, got {T}")
    width = _plan.resolve_panel_width(plan, N, k, T, override=w)
    source = GhostSource(int(seed) & 0xFFFFFFFFFFFFFFFF, int(N), int(k), width)
    panels = []
    for j in range(_plan.cld(int(k), width)):
        c0, c1 = j * width, min((j + 1) * width, int(k))
        # 1-based column indices in the full matrix, as the stream is keyed.
        panels.append(Panel(None, GHOST, GhostHome(source, (c0 + 1, c1)), c1 - c0))
    return PanelMatrix(N, k, plan=plan, w=width, dtype=T, _panels=panels)
