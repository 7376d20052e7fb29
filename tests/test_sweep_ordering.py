"""Ordering discipline of the sweep pipeline, on the CPU.

* The jittered backend (tests/fakes.py) runs every queue on its own worker
  thread with a random sleep before each operation, so the only ordering
  between uploads, kernels and write-backs is the events the sweep waits on,
  as on the device. gram, cholqr2 (both schedules), panelmul, project (both
  schedules), scale and dot are run 50 times at nbuffers 1 and 2 and must
  give the bits of an unjittered nbuffers = 1 run (the reference's jittered
  repetitions, test/setup.jl:105-126, src/device_api.jl:193-256).
* A planted ordering mistake (every wait_event dropped) must be caught.
* No synchronization call sits inside a loop of the host code, bar the
  synchronization helpers themselves (test/test_discipline.jl:1-54).
"""
import ast
import os

import numpy as np
import pytest

import funicular_hip as F
from fakes import JitteryBackend

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACKAGE = os.path.join(ROOT, "funicular.jl_amd", "funicular_hip")

N, K, W = 96, 12, 4  # three panels, three row blocks of 32


def _matrix(seed, N=N, k=K):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k)))


def _plan(backend, **kw):
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20, **kw)


def _operations():
    """(name, fn(plan, backend, nb) -> array) for every panel operation that
    runs a sweep."""
    A, B = _matrix(1), _matrix(2)
    G = _matrix(3, N, N) / np.sqrt(N)

    def gram(plan, backend, nb):
        return F.gram(F.PanelMatrix.from_array(A, plan=plan, w=W), nbuffers=nb)

    def gram2(plan, backend, nb):
        return F.gram(F.PanelMatrix.from_array(A, plan=plan, w=W), F.PanelMatrix.from_array(B, plan=plan, w=W),
                      nbuffers=nb)

    def cholqr2(schedule):
        def run(plan, backend, nb):
            Y = F.PanelMatrix.from_array(A, plan=plan, w=W)
            R = F.cholqr2(Y, nbuffers=nb, schedule=schedule)
            return np.concatenate([F.to_matrix(Y), R])
        return run

    def panelmul(plan, backend, nb):
        X = F.PanelMatrix.from_array(A, plan=plan, w=W)
        Y = F.PanelMatrix(N, K, plan=plan, w=W)
        F.panelmul(Y, F.DenseOperator(G, backend), X, nbuffers=nb)
        return F.to_matrix(Y)

    def project(schedule, group=None):
        def run(plan, backend, nb):
            Q = F.PanelMatrix.from_array(A, plan=plan, w=W)
            return F.project(Q, F.DenseOperator(G, backend), nbuffers=nb, schedule=schedule, group=group)
        return run

    def scale_dot(plan, backend, nb):
        X = F.PanelMatrix.from_array(A, plan=plan, w=W)
        Y = F.PanelMatrix.from_array(B, plan=plan, w=W)
        F.scale(Y, 0.5 - 0.25j, nbuffers=nb)
        F.axpy(2.0, X, Y, nbuffers=nb)
        d = F.dot(X, Y, nbuffers=nb)
        return np.concatenate([F.to_matrix(Y).ravel(), [d]])

    def ghost_host(plan, backend, nb):
        O = F.GhostPanels(F.randn_column, np.complex128, N, K, plan=plan, seed=9, w=W)
        Y = F.PanelMatrix(N, K, plan=plan, w=W)
        F.panelmul(Y, F.DenseOperator(G, backend), O, nbuffers=nb)
        return np.concatenate([F.to_matrix(Y), F.gram(O, nbuffers=nb)])

    return [("ghost_host", ghost_host), ("gram", gram), ("gram_xy", gram2), ("cholqr2_reference", cholqr2("reference")),
            ("cholqr2_fused", cholqr2("fused")), ("panelmul", panelmul), ("project_reference", project("reference")),
            ("project_runs", project("fused", group=2)), ("scale_axpy_dot", scale_dot)]


@pytest.mark.parametrize("name,op", _operations(), ids=[n for n, _ in _operations()])
def test_jittered_sweeps_give_the_unjittered_bits(name, op):
    calm = JitteryBackend(jitter=0.0)
    expected = op(_plan(calm), calm, 1)
    for rep in range(50):
        backend = JitteryBackend(jitter=2e-4, seed=rep)
        got = op(_plan(backend), backend, 1 + rep % 2)
        assert got.tobytes() == expected.tobytes(), (name, rep)


def test_a_dropped_wait_is_caught():
    # With every wait_event ignored, the compute queue can read a staging
    # buffer before its upload lands (or the upload of a later step can
    # overwrite it): some repetition must come out different.
    A = _matrix(1)
    calm = JitteryBackend(jitter=0.0)
    plan = _plan(calm)
    expected = F.gram(F.PanelMatrix.from_array(A, plan=plan, w=W), nbuffers=2)
    differs = 0
    for rep in range(20):
        backend = JitteryBackend(jitter=1e-3, seed=rep, drop_waits=True)
        plan = _plan(backend)
        got = F.gram(F.PanelMatrix.from_array(A, plan=plan, w=W), nbuffers=2)
        differs += got.tobytes() != expected.tobytes()
    assert differs > 0


SYNCHRONIZATION = {"sync_all", "sync_queue", "sync_event", "sync_quietly", "fhip_stream_sync", "fhip_event_sync",
                   "fhip_device_sync"}


def _called_name(node):
    f = node.func
    if isinstance(f, ast.Attribute):
        return f.attr
    if isinstance(f, ast.Name):
        return f.id
    return None


def syncs_inside_loops(tree):
    """Names of synchronization calls that sit inside a for or while loop,
    outside functions whose own name starts with "sync"."""
    found = []

    def visit(node, inloop):
        if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef)):
            if node.name.startswith("sync"):
                return
            for child in node.body:
                visit(child, False)
            return
        if isinstance(node, ast.Call) and inloop:
            name = _called_name(node)
            if name in SYNCHRONIZATION or (name == "call" and node.args and isinstance(node.args[0], ast.Constant)
                                           and node.args[0].value in SYNCHRONIZATION):
                found.append(name if name != "call" else node.args[0].value)
        nested = inloop or isinstance(node, (ast.For, ast.While, ast.AsyncFor))
        for child in ast.iter_child_nodes(node):
            visit(child, nested)

    visit(tree, False)
    return found


def test_no_synchronization_inside_a_loop():
    for fname in sorted(os.listdir(PACKAGE)):
        if fname.endswith(".py"):
            path = os.path.join(PACKAGE, fname)
            assert syncs_inside_loops(ast.parse(open(path).read(), path)) == [], fname


def test_the_loop_detector_has_teeth():
    offender = ast.parse("def sweep():\n    for j in range(3):\n        backend.sync_queue(q)\n")
    assert syncs_inside_loops(offender) == ["sync_queue"]
    allowed = ast.parse("def sweep():\n    for j in range(3):\n        pass\n    backend.sync_queue(q)\n")
    assert syncs_inside_loops(allowed) == []
    whitelisted = ast.parse("def sync_them(qs):\n    for q in qs:\n        backend.sync_queue(q)\n")
    assert syncs_inside_loops(whitelisted) == []
    raw = ast.parse("def f():\n    while True:\n        call('fhip_stream_sync', s)\n")
    assert syncs_inside_loops(raw) == ["fhip_stream_sync"]
