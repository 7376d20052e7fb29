"""The reference's panel/tier schedule on the HIP backend, step for step and
byte for byte: the sweeps each operation runs (their traversals, steps,
labels and operand modes) equal the oracle's restatement of the reference's
(oracle/schedule.py, itself pinned to the reference's own statements), and
the library's copy counters equal the bytes and copy calls that schedule
issues. The shape is configuration 1's (2^20 x 64 in panels of 8), host
resident, at nbuffers 1 and 2; a ragged shape and the shifted third pass
follow."""
import numpy as np
import pytest

import funicular_hip as F
from funicular_hip import _lib
from funicular_hip.executor import trace_sweeps
from oracle import operators as OO
from oracle import panelblas as OB
from oracle import schedule as OS

pytestmark = pytest.mark.gpu

C16 = np.complex128
COUNTED = ("h2d_bytes", "d2h_bytes", "copies", "d2d_bytes")


def _plan(backend, **kw):
    kw.setdefault("device_budget", 4 << 30)
    kw.setdefault("host_budget", 4 << 30)
    return F.ResidencyPlan(backend=backend, **kw)


def _normal(rng, N, k):
    return np.asfortranarray((rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k))) / np.sqrt(2))


def _counted(fn):
    """Runs fn, returning its result, the sweeps it issued and the change in
    the library's copy counters."""
    before = _lib.counters()
    with trace_sweeps() as records:
        out = fn()
    after = _lib.counters()
    return out, records, {name: after[name] - before[name] for name in COUNTED}


def _steps(records):
    return [{key: r[key] for key in ("traversal", "steps", "modes")} for r in records]


def _expected(events, N, k, w):
    return dict(OS.transfers(events, N, k, w, 16), d2d_bytes=0)


@pytest.mark.parametrize("nb", (1, 2))
def test_cholqr2_runs_the_reference_schedule_at_configuration_1(backend, nb):
    N, k, w = 1 << 20, 64, 8
    print(f"reference schedule cholqr2 N={N} k={k} w={w} nb={nb}", flush=True)
    A = _normal(np.random.default_rng(4100), N, k)
    plan = _plan(backend)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R, records, moved = _counted(lambda: F.cholqr2(Y, nbuffers=nb))
    events = OS.cholqr2(N, k, w)
    assert _steps(records) == OS.sweeps(events)
    assert all(r["staged"] == (True,) and r["nbuffers"] == nb for r in records)
    assert moved == _expected(events, N, k, w)
    Q = F.to_matrix(Y)
    Qo, Ro = OB.cholqr2(A, w)
    assert OB.orthogonality_error(Q) <= 1e-13
    assert np.linalg.norm(Q @ R - A) / np.linalg.norm(A) <= 1e-13
    assert np.linalg.norm(R - Ro) <= 1e-12 * np.linalg.norm(Ro)
    assert np.linalg.norm(Q - Qo) <= 1e-12 * np.linalg.norm(Qo)


def test_cholqr2_schedule_gives_the_same_bits_at_nbuffers_1_and_2(backend):
    N, k, w = 200000, 64, 8
    A = _normal(np.random.default_rng(4101), N, k)
    out = []
    for nb in (1, 2):
        plan = _plan(backend)
        Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
        R = F.cholqr2(Y, nbuffers=nb)
        out.append((R, F.to_matrix(Y)))
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()


@pytest.mark.parametrize("nb", (1, 2))
def test_gram_project_and_panelmul_run_the_reference_schedule(backend, nb):
    N, k, w = 1 << 20, 64, 8
    print(f"reference schedule gram/project/panelmul nb={nb}", flush=True)
    rng = np.random.default_rng(4102)
    A, B = _normal(rng, N, k), _normal(rng, N, k)
    plan = _plan(backend)
    X = F.PanelMatrix.from_array(A, plan=plan, w=w)
    Z = F.PanelMatrix.from_array(B, plan=plan, w=w)
    d, o = 2 + 0.1j, -1.0
    G = F.Tridiagonal1D(N, d, o, backend=backend)

    Gx, records, moved = _counted(lambda: F.gram(X, nbuffers=nb))
    ev = OS.gram(N, k, w)
    assert _steps(records) == OS.sweeps(ev) and moved == _expected(ev, N, k, w)
    assert OB.isapprox(Gx, OB.gram(A, w=w), OB.blastol(C16, N, k))

    Gxz, records, moved = _counted(lambda: F.gram(X, Z, nbuffers=nb))
    ev = OS.gram(N, k, w, two_operands=True)
    assert _steps(records) == OS.sweeps(ev) and moved == _expected(ev, N, k, w)
    assert OB.isapprox(Gxz, OB.gram(A, B, w=w), OB.blastol(C16, N, k))

    S, records, moved = _counted(lambda: F.project(X, G, nbuffers=nb))
    ev = OS.project(N, k, w)
    assert _steps(records) == OS.sweeps(ev) and moved == _expected(ev, N, k, w)
    So = OB.project(A, lambda V: OO.tridiag_apply(V, d, o), w)
    assert OB.isapprox(S, So, OB.blastol(C16, N, k))

    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    _, records, moved = _counted(lambda: F.panelmul(Y, G, X, nbuffers=nb))
    ev = OS.panelmul(N, k, w)
    assert _steps(records) == OS.sweeps(ev) and moved == _expected(ev, N, k, w)
    # The stencil's operation order is the kernel's own (fused multiply-adds),
    # so it matches the oracle to rounding, as in the kernel tests.
    Yo = OB.panelmul(lambda V: OO.tridiag_apply(V, d, o), A, w)
    assert np.allclose(F.to_matrix(Y), Yo, rtol=1e-14, atol=1e-14)


def test_reference_schedule_on_a_ragged_shape(backend):
    # A last panel of 2 columns and a last row block shorter than the rest.
    N, k, w = 100003, 23, 7
    A = _normal(np.random.default_rng(4103), N, k)
    plan = _plan(backend, device_budget=512 << 20, host_budget=512 << 20)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R, records, moved = _counted(lambda: F.cholqr2(Y))
    events = OS.cholqr2(N, k, w)
    assert _steps(records) == OS.sweeps(events)
    assert moved == _expected(events, N, k, w)
    assert OB.orthogonality_error(F.to_matrix(Y)) <= 1e-13
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, backend=backend)
    _, records, moved = _counted(lambda: F.project(Y, G))
    ev = OS.project(N, k, w)
    assert _steps(records) == OS.sweeps(ev) and moved == _expected(ev, N, k, w)


def test_a_shifted_pass_adds_a_third_reference_pass(backend):
    # A zero column makes the Gram singular, the plain Cholesky fails, the
    # shift fires and a third pass runs (src/panelblas.jl:130-139, 148-155;
    # the zero-column case of test/test_panelblas.jl:327-345). The zero
    # column stays zero and the others come out orthonormal.
    N, k, w = 50000, 16, 4
    A = _normal(np.random.default_rng(4104), N, k)
    A[:, 5] = 0
    plan = _plan(backend, device_budget=256 << 20, host_budget=256 << 20)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R, records, moved = _counted(lambda: F.cholqr2(Y))
    events = OS.cholqr2(N, k, w, shifted=True)
    assert _steps(records) == OS.sweeps(events)
    assert moved == _expected(events, N, k, w)
    Q = F.to_matrix(Y)
    assert np.all(Q[:, 5] == 0)
    # Each shifted pass leaves Q^H Q = I - s (G + s I)^-1, about 1e-8 off here:
    # the shift costs orthogonality, as the reference's docstring warns.
    keep = [c for c in range(k) if c != 5]
    assert OB.orthogonality_error(Q[:, keep]) <= 1e-6
    Qo, Ro = OB.cholqr2(A, w)
    assert np.linalg.norm(Q - Qo) <= 1e-10 * np.linalg.norm(Qo)
    assert np.linalg.norm(R - Ro) <= 1e-10 * np.linalg.norm(Ro)


def test_reference_schedule_on_a_device_resident_plan(backend):
    # Panels in HBM: the same row-block sweeps run on views of the matrix, so
    # only the k x k copies cross the link (the Gram down and R up per pass).
    N, k, w = 1 << 20, 64, 8
    A = _normal(np.random.default_rng(4105), N, k)
    plan = _plan(backend, device_resident=True)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    R, records, moved = _counted(lambda: F.cholqr2(Y))
    events = OS.cholqr2(N, k, w)
    assert _steps(records) == OS.sweeps(events)
    assert all(r["staged"] == (False,) for r in records)
    assert moved == {"h2d_bytes": 2 * k * k * 16, "d2h_bytes": 2 * k * k * 16, "copies": 4, "d2d_bytes": 0}
    # The same bits as the host-resident run: the kernels see the same blocks.
    hp = _plan(backend)
    Yh = F.PanelMatrix.from_array(A, plan=hp, w=w)
    Rh = F.cholqr2(Yh)
    assert R.tobytes() == Rh.tobytes()
    assert F.to_matrix(Y).tobytes() == F.to_matrix(Yh).tobytes()


def test_the_fused_schedule_is_an_opt_in(backend):
    # The plan's schedule="fused" (or the keyword) runs three uploads of Y
    # and keeps the Cholesky on the device; the default runs four.
    N, k, w = 200000, 64, 8
    A = _normal(np.random.default_rng(4106), N, k)
    ups = {}
    Rs = {}
    for schedule in ("reference", "fused"):
        plan = _plan(backend, schedule=schedule)
        Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
        Rs[schedule], _, moved = _counted(lambda: F.cholqr2(Y))
        ups[schedule] = moved["h2d_bytes"] // (N * k * 16)
    assert ups == {"reference": 4, "fused": 3}
    assert np.linalg.norm(Rs["fused"] - Rs["reference"]) <= 1e-13 * np.linalg.norm(Rs["reference"])


@pytest.mark.parametrize("schedule", ("reference", "fused"))
def test_no_host_wait_inside_a_sweeps_issue_loop(backend, schedule):
    # The library counts every blocking wait (fhip_get_counters host_syncs).
    # Read at each step a sweep issues, it must not move within a sweep: the
    # issue loop queues work and waits only once, after its last step
    # (test/test_discipline.jl:13-45; src/executor.jl:193-226).
    N, k, w = 300000, 64, 8
    A = _normal(np.random.default_rng(4107), N, k)
    plan = _plan(backend, schedule=schedule)
    Y = F.PanelMatrix.from_array(A, plan=plan, w=w)
    seen = []
    originals = {name: getattr(backend, name) for name in ("gram_accumulate", "rdiv_upper", "rdiv_upper_folded")}

    def watched(name):
        def call(*args, **kwargs):
            seen.append((len(records), _lib.counters()["host_syncs"]))
            return originals[name](*args, **kwargs)
        return call
    try:
        for name in originals:
            setattr(backend, name, watched(name))
        with trace_sweeps() as records:
            F.cholqr2(Y)
    finally:
        for name, fn in originals.items():
            setattr(backend, name, fn)
    by_sweep = {}
    for sweep_index, syncs in seen:
        by_sweep.setdefault(sweep_index, set()).add(syncs)
    assert len(by_sweep) >= 3
    assert all(len(values) == 1 for values in by_sweep.values()), by_sweep
