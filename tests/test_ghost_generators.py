"""GhostPanels with a host column generator, the reference's form
GhostPanels(f!, T, N, k) (src/ghosts.jl:53-122), on the CPU through the
jittered backend: the generator sees 1-based columns, the values do not
depend on the panel width or the visiting order, the seed changes them, and
a sweep reads them as an ordinary host-resident matrix
(test/test_ghosts.jl:1-56, 88-100 are the model)."""
import numpy as np
import pytest

import funicular_hip as F
from fakes import JitteryBackend


def _plan(backend=None, **kw):
    backend = JitteryBackend(jitter=0.0) if backend is None else backend
    return F.ResidencyPlan(backend=backend, device_budget=64 << 20, host_budget=64 << 20, **kw)


def test_the_generator_sees_one_based_columns():
    seen = []

    def f(dst, rng, col):
        seen.append(col)
        dst[...] = col
    plan = _plan()
    O = F.GhostPanels(f, np.float64, 7, 10, plan=plan, w=4)
    A = F.to_matrix(O)
    assert np.array_equal(A, np.tile(np.arange(1, 11, dtype=float), (7, 1)))
    assert sorted(set(seen)) == list(range(1, 11))


@pytest.mark.parametrize("T", [np.complex128, np.float64, np.complex64])
def test_width_independent_and_seed_sensitive(T):
    plan = _plan()
    ref = F.to_matrix(F.GhostPanels(F.randn_column, T, 300, 13, plan=plan, seed=0xDEADBEEF, w=13))
    for w in (1, 4, 7):
        got = F.to_matrix(F.GhostPanels(T, 300, 13, plan=plan, seed=0xDEADBEEF, w=w, generator=F.randn_column))
        assert np.array_equal(got, ref)
    other = F.to_matrix(F.GhostPanels(F.randn_column, T, 300, 13, plan=plan, seed=0xDEADBEEF + 1, w=4))
    assert not np.array_equal(other, ref)
    # Standard normal entries, complex parts of variance 1/2 each.
    v = np.var(ref.real) + (np.var(ref.imag) if np.iscomplexobj(ref) else 0.0)
    assert 0.8 < v < 1.2


def test_sweeps_read_host_generated_panels_in_any_order():
    # Row sweeps stage every panel; panel sweeps one at a time, regenerating
    # whatever the host budget let go.
    backend = JitteryBackend(jitter=1e-4, seed=3)
    plan = _plan(backend)
    O = F.GhostPanels(F.randn_column, np.complex128, 200, 12, plan=plan, seed=11, w=4)
    A = F.to_matrix(O)
    assert np.allclose(F.gram(O), A.conj().T @ A, rtol=1e-13, atol=1e-12)
    Y = F.PanelMatrix(200, 12, plan=plan, w=4)
    F.panelmul(Y, F.DenseOperator(np.eye(200), backend), O)
    assert np.array_equal(F.to_matrix(Y), A)


def test_a_ghost_with_a_host_generator_refuses_writes_and_bad_arguments():
    plan = _plan()
    O = F.GhostPanels(F.randn_column, np.complex128, 50, 6, plan=plan, w=3)
    with pytest.raises(F.ArgumentError):
        F.scale(O, 2.0)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(F.randn_column, np.complex128, 50, 6, plan=plan, generator=F.randn_column)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(np.complex128, 50, 6, plan=plan, generator=3)
    with pytest.raises(F.ArgumentError):
        F.GhostPanels(np.complex128, 0, 6, plan=plan)


def test_the_default_generator_is_randn_column_and_the_device_stream_is_named():
    # GhostPanels(T, N, k) takes the host generator, as the reference's takes
    # randn! (src/ghosts.jl:53-63); the device stream is chosen by name.
    plan = _plan()
    default = F.to_matrix(F.GhostPanels(np.complex128, 120, 9, plan=plan, seed=5, w=4))
    named = F.to_matrix(F.GhostPanels(F.randn_column, np.complex128, 120, 9, plan=plan, seed=5, w=4))
    assert np.array_equal(default, named)
    device = F.GhostPanels(F.device_randn, np.complex128, 120, 9, plan=plan, seed=5, w=4)
    assert all(p.home.source.generator is None for p in device.panels)
    assert repr(F.device_randn) == "device_randn"
