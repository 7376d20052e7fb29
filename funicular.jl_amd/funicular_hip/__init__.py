"""funicular_hip: Funicular.jl's panel algebra on an MI355X, behind the
reference's device contract.

The public names mirror Funicular's exports (src/Funicular.jl:18-20); Julia's
mutating `f!` is `f` here (cholqr2 overwrites Y as cholqr2! does), and panel
and column numbers start at 0. `hip_backend()` is the device backend; it is
the only compute path, and importing this package fails loudly when
libfunicular_hip.so has not been built.
"""
from . import _lib

_lib.load()

from .device_api import (ArgumentError, DeviceArray, DimensionMismatch, HIPBackend, hip_backend,  # noqa: E402
                         human_readable_bytes, slab_matrix)
from .adjoint import rrule_cholqr2, rrule_project  # noqa: E402
from .executor import (PanelSteps, RowSteps, foreachpanel, readpanels, rowtraversal, sweep,  # noqa: E402
                       updatepanels, writepanels)
from .ghosts import GhostPanels, device_randn, randn_column  # noqa: E402
from .io import disk_reads, disk_writes, load, save  # noqa: E402
from .operator import (DenseOperator, Helmholtz3D, Tridiagonal1D, check_operator, ishermitian_op,  # noqa: E402
                       panel_capable, workspace_bytes)
from .panelblas import (axpy, cholesky_upper, cholqr2, cholqr2_mixed, cholqr2_panelwise, dot, gram,  # noqa: E402
                        gram_panelwise, norm, panelmul, project, rdiv_panelwise, rdiv_rows, rightmul, scale)
from .panels import (PanelMatrix, copycols, copyto, evict, free, npanels, panelrange, panelwidth,  # noqa: E402
                     prefetch, similar, to_matrix)
from .plan import (ResidencyPlan, choose_panel_width, device_bytes_allocated, resolve_panel_width,  # noqa: E402
                   row_block_height)

__all__ = [
    "GhostPanels", "device_randn", "randn_column", "PanelMatrix", "ResidencyPlan", "check_operator", "cholqr2", "copycols", "foreachpanel",
    "gram", "npanels", "panelmul", "panelrange", "panelwidth", "project", "rightmul", "hip_backend",
    "rrule_cholqr2", "rrule_project", "save", "load", "disk_reads", "disk_writes", "scale", "axpy", "dot", "norm",
    "cholqr2_panelwise", "gram_panelwise", "rdiv_panelwise",
]
