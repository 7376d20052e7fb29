"""The disk tier on the host side (test/test_disk.jl is the model): panels
spill to the matrix's scratch file when the host budget runs out and come back
bit for bit, narrowed storage, eviction and prefetch, save and load, read-only
stores, and the file format itself. No device is needed: every copy here is
host to host or host to file."""
import json
import os

import numpy as np
import pytest

from fakes import HostOnlyBackend
import funicular_hip as F
from funicular_hip import io as FIO
from funicular_hip import plan as P
from funicular_hip.panels import evict, free, materialize, ondisk, prefetch, to_matrix

C16 = np.complex128
ELTYPES = (np.complex128, np.complex64, np.float64, np.float32)


def randmatrix(rng, N, k, T):
    A = rng.random((N, k)) - 0.5
    if np.issubdtype(T, np.complexfloating):
        A = A + 1j * (rng.random((N, k)) - 0.5)
    return np.asfortranarray(A.astype(T))


def panelbudget(n, N, w, T):
    """Host budget for n panels of N x w (test/setup.jl's panelbudget)."""
    return n * P.align_up(P.panel_bytes(N, w, np.dtype(T)), P.SLAB_ALIGN)


def make_plan(tmp_path, **kw):
    kw.setdefault("device_budget", 1 << 30)
    return F.ResidencyPlan(backend=HostOnlyBackend(), scratch_dir=str(tmp_path), **kw)


@pytest.mark.parametrize("T", ELTYPES)
@pytest.mark.parametrize("w", (8, 7, 1))
def test_three_tier_round_trip(tmp_path, T, w):
    rng = np.random.default_rng(w)
    N, k = 12, 32
    A = randmatrix(rng, N, k, T)
    plan = make_plan(tmp_path, panel_width=w, host_budget=panelbudget(2, N, w, T))
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    assert F.npanels(pm) > 2
    assert sum(ondisk(p) for p in pm.panels) == F.npanels(pm) - 2
    assert np.array_equal(to_matrix(pm), A)
    assert F.disk_reads(pm) > 0 and F.disk_writes(pm) > 0
    assert len(os.listdir(tmp_path)) == 1
    free(pm)
    assert os.listdir(tmp_path) == []


def test_narrowed_storage_through_the_disk_tier(tmp_path):
    rng = np.random.default_rng(1)
    N, k, w = 9, 20, 4
    A = randmatrix(rng, N, k, C16)
    plan = make_plan(tmp_path, panel_width=w, host_eltype=np.complex64,
                     host_budget=panelbudget(2, N, w, np.complex64))
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    assert pm.store.stored == np.dtype(np.complex64)
    assert pm.store.computed == np.dtype(C16)
    assert np.array_equal(to_matrix(pm), A.astype(np.complex64).astype(C16))
    assert F.disk_reads(pm) > 0


def test_eviction_and_prefetch(tmp_path):
    rng = np.random.default_rng(2)
    N, k, w = 8, 12, 4
    A = randmatrix(rng, N, k, C16)
    plan = make_plan(tmp_path, panel_width=w, host_budget=panelbudget(3, N, w, C16))
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    assert not any(ondisk(p) for p in pm.panels)
    evict(pm.panels[1], plan)
    assert ondisk(pm.panels[1])
    assert evict(pm.panels[1], plan) is None
    with pytest.raises(F.ArgumentError):
        materialize(plan.backend.alloc_device(C16, (N, w)), pm.panels[1], plan)
    prefetch(pm, 1)
    assert not ondisk(pm.panels[1])
    assert np.array_equal(F.panels.panelstorage(pm.panels[1]), A[:, 4:8])
    assert np.array_equal(to_matrix(pm), A)
    # A panel pinned by a running sweep cannot be evicted.
    P.pin(plan, pm.panels[0])
    with pytest.raises(F.ArgumentError):
        evict(pm.panels[0], plan)
    P.unpin(plan, pm.panels[0])


def test_evicting_without_a_colder_tier_is_an_error():
    plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20)
    pm = F.PanelMatrix.from_array(np.zeros((8, 4), C16), plan=plan, w=2)
    with pytest.raises(F.ArgumentError, match="scratch_dir"):
        evict(pm.panels[0], plan)


def test_clean_panels_are_not_written_twice(tmp_path):
    rng = np.random.default_rng(3)
    N, k, w = 10, 16, 4
    A = randmatrix(rng, N, k, C16)
    plan = make_plan(tmp_path, panel_width=w, host_budget=panelbudget(4, N, w, C16))
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    for p in pm.panels:
        evict(p, plan)
    writes = F.disk_writes(pm)
    prefetch(pm, 0)
    evict(pm.panels[0], plan)
    assert F.disk_writes(pm) == writes
    F.copycols(pm, range(0, 1), np.ones((N, 1), C16))
    evict(pm.panels[0], plan)
    assert F.disk_writes(pm) == writes + 1
    B = A.copy()
    B[:, 0] = 1
    assert np.array_equal(to_matrix(pm), B)


@pytest.mark.parametrize("T", ELTYPES)
def test_save_and_load(tmp_path, T):
    rng = np.random.default_rng(4)
    N, k, w = 13, 10, 3
    A = randmatrix(rng, N, k, T)
    plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20)
    pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
    path = str(tmp_path / "m.panels")
    assert F.save(pm, path) == path
    assert np.array_equal(to_matrix(pm), A)
    loaded = F.load(path, plan=plan)
    assert (loaded.N, loaded.k, loaded.w, loaded.dtype) == (N, k, w, np.dtype(T))
    assert all(ondisk(p) for p in loaded.panels)
    assert np.array_equal(to_matrix(loaded), A)
    assert F.disk_reads(loaded) == F.npanels(loaded)
    # The file is the loaded matrix's cold tier: a change goes back into it.
    F.copycols(loaded, range(2, 3), np.full((N, 1), 7, T))
    for p in loaded.panels:
        evict(p, plan)
    free(loaded)
    again = F.load(path, plan=plan)
    B = A.copy()
    B[:, 2] = 7
    assert np.array_equal(to_matrix(again), B)
    free(again)
    assert os.path.exists(path)  # a loaded file is not scratch


def test_file_format(tmp_path):
    rng = np.random.default_rng(5)
    N, k, w = 7, 9, 4
    A = randmatrix(rng, N, k, C16)
    plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20)
    path = str(tmp_path / "m.panels")
    F.save(F.PanelMatrix.from_array(A, plan=plan, w=w), path)
    raw = open(path, "rb").read()
    assert raw.startswith(FIO.MAGIC)
    header = json.loads(raw[len(FIO.MAGIC):FIO.HEADER_BYTES].rstrip(b"\0").decode())
    assert header == {"funicular_format": 1, "dataset": "panels", "N": N, "k": k, "panel_width": w,
                      "storage_eltype": "ComplexF64", "compute_eltype": "ComplexF64", "order": "column-major"}
    data = np.frombuffer(raw, dtype=C16, offset=FIO.HEADER_BYTES)
    assert np.array_equal(data, A.ravel(order="F"))
    assert len(raw) == FIO.HEADER_BYTES + N * k * 16


def test_readonly_load_refuses_writes(tmp_path):
    rng = np.random.default_rng(6)
    N, k, w = 6, 8, 4
    A = randmatrix(rng, N, k, C16)
    plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20)
    path = str(tmp_path / "m.panels")
    F.save(F.PanelMatrix.from_array(A, plan=plan, w=w), path)
    ro = F.load(path, plan=plan, readonly=True)
    assert np.array_equal(to_matrix(ro), A)
    F.copycols(ro, range(0, 1), np.zeros((N, 1), C16))
    with pytest.raises(F.ArgumentError, match="read only"):
        evict(ro.panels[0], plan)
    free(ro)


def test_load_errors(tmp_path):
    plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20)
    with pytest.raises(F.ArgumentError, match="no file"):
        F.load(str(tmp_path / "missing.panels"), plan=plan)
    junk = tmp_path / "junk.panels"
    junk.write_bytes(b"not a panel file" * 300)
    with pytest.raises(F.ArgumentError, match="not one Funicular wrote"):
        F.load(str(junk), plan=plan)
    path = str(tmp_path / "m.panels")
    F.save(F.PanelMatrix.from_array(np.ones((5, 4), C16), plan=plan, w=2), path)
    # A plan whose cold tiers hold another eltype cannot take the file.
    narrow = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, host_budget=1 << 20,
                             host_eltype=np.complex64)
    with pytest.raises(F.ArgumentError, match="stores"):
        F.load(path, plan=narrow)
    # A truncated file is refused before any panel is read.
    with open(path, "r+b") as f:
        f.truncate(FIO.HEADER_BYTES + 10)
    with pytest.raises(F.ArgumentError, match="short"):
        F.load(path, plan=plan)


def test_closed_store_and_scratch_cleanup(tmp_path):
    N, k, w = 6, 8, 2
    plan = make_plan(tmp_path, panel_width=w, host_budget=panelbudget(2, N, w, C16))
    pm = F.PanelMatrix.from_array(np.ones((N, k), C16), plan=plan, w=w)
    cold = [p for p in pm.panels if ondisk(p)]
    assert cold
    store = pm.store
    FIO.close_store(store)
    assert not os.path.exists(store.path)
    with pytest.raises(F.ArgumentError, match="closed"):
        FIO.read_panel(cold[0].home, np.empty((N, w), C16, order="F"))
    assert FIO.close_store(store) is None


def test_host_full_without_spillable_panels(tmp_path):
    N, k, w = 8, 16, 4
    plan = make_plan(tmp_path, panel_width=w, host_budget=panelbudget(1, N, w, C16))
    pm = F.PanelMatrix.from_array(np.ones((N, k), C16), plan=plan, w=w)
    for p in pm.panels:
        if not ondisk(p):
            P.pin(plan, p)
    with pytest.raises(F.ArgumentError, match="no panel of that size is free to spill"):
        prefetch(pm, next(j for j, p in enumerate(pm.panels) if ondisk(p)))


def test_disk_staging_under_jitter(tmp_path):
    # test/test_disk.jl:110-127: a write sweep whose panels cycle through the
    # scratch file, on jittered queues
    from fakes import JitteryBackend
    rng = np.random.default_rng(40)
    N, k, w = 12, 20, 4
    A = randmatrix(rng, N, k, C16)
    for rep in range(10):
        backend = JitteryBackend(jitter=2e-4, seed=rep)
        d = tmp_path / f"rep{rep}"
        d.mkdir()
        plan = F.ResidencyPlan(backend=backend, device_budget=1 << 30, scratch_dir=str(d), panel_width=w,
                               host_budget=panelbudget(4, N, w, C16))
        pm = F.PanelMatrix.from_array(A, plan=plan, w=w)
        F.foreachpanel(lambda j, block: backend.scale(block, 2.0), pm)
        assert np.array_equal(to_matrix(pm), 2 * A)
        assert all(p.pins == 0 for p in pm.panels)
        free(pm)


def test_concurrent_readers_of_one_store(tmp_path):
    # test/test_disk.jl:238-259: four threads load one file read-only, each
    # with room for three of its eight panels, and collect it five times
    import threading
    rng = np.random.default_rng(41)
    N, k, w = 9, 24, 3
    A = randmatrix(rng, N, k, C16)
    path = str(tmp_path / "shared.fpan")
    F.save(F.PanelMatrix.from_array(A, plan=F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30,
                                                             host_budget=1 << 20), w=w), path)
    results = [None] * 4

    def reader(i):
        plan = F.ResidencyPlan(backend=HostOnlyBackend(), device_budget=1 << 30, panel_width=w,
                               host_budget=panelbudget(3, N, w, C16))
        pm = F.load(path, plan=plan, readonly=True)
        try:
            results[i] = all(np.array_equal(to_matrix(pm), A) for _ in range(5))
        finally:
            free(pm)

    threads = [threading.Thread(target=reader, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert results == [True] * 4
