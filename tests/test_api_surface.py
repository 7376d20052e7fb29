"""The public names, after test/test_api_surface.jl: the reference's exports
(src/Funicular.jl:18-20) all have a counterpart here, each documented, and
the host package imports no vendor or framework package (the compute path is
libfunicular_hip.so alone)."""
import ast
import os

import funicular_hip as F

PACKAGE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "funicular.jl_amd",
                       "funicular_hip")

# The reference's exports and the name each has here: Julia's mutating f! is
# f, and the CPU backend's place is taken by the HIP backend.
REFERENCE_EXPORTS = {
    "CPUBackend": "hip_backend", "GhostPanels": "GhostPanels", "PanelMatrix": "PanelMatrix",
    "ResidencyPlan": "ResidencyPlan", "check_operator": "check_operator", "cholqr2!": "cholqr2",
    "copycols!": "copycols", "foreachpanel": "foreachpanel", "gram": "gram", "load": "load",
    "npanels": "npanels", "panelmul!": "panelmul", "panelrange": "panelrange", "panelwidth": "panelwidth",
    "project": "project", "rightmul!": "rightmul", "save": "save", "scale!": "scale",
}


def test_every_reference_export_has_a_counterpart():
    for julia, here in REFERENCE_EXPORTS.items():
        assert hasattr(F, here), julia
        assert here in F.__all__, julia


def test_exports_are_documented():
    for name in F.__all__:
        obj = getattr(F, name)
        assert (getattr(obj, "__doc__", None) or "").strip(), name


def test_the_package_imports_no_vendor_package():
    forbidden = {"torch", "cupy", "jax", "triton", "h5py", "numba", "pycuda"}
    for root, _, files in os.walk(PACKAGE):
        for f in files:
            if not f.endswith(".py"):
                continue
            tree = ast.parse(open(os.path.join(root, f)).read())
            for node in ast.walk(tree):
                if isinstance(node, ast.Import):
                    mods = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom) and node.module and node.level == 0:
                    mods = [node.module]
                else:
                    continue
                for m in mods:
                    assert m.split(".")[0] not in forbidden, (f, m)
