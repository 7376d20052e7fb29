"""The sweep pipeline (src/executor.jl:1-267).

A sweep visits panel matrices step by step. Each step's read operands are
copied into device staging buffers on the upload queue, the panel function runs
on the compute queue, and write operands are copied back on the writeback
queue; only events order the three. With nbuffers=2 the uploads run one
step ahead of the compute, and a slot's buffer is refilled only once the step that used
it nbuffers steps before is done with it. A producer thread stays FEED_DEPTH
steps ahead, bringing each step's panels to where the copies can read them.
The issue loop allocates nothing (every buffer is the plan's) and never waits;
the host blocks once, after the last step.

Device-resident operands (plans with device_resident=True) need no staging:
their steps are views of the matrix's own HBM allocation.
"""
import contextlib
import queue as _queue
import threading

from . import plan as _plan
from .device_api import ArgumentError
from .panels import (assert_same_panels, assert_same_rows, assert_writable, checkout_staging_buffers,
                     materialize, materialize_rows, npanels, panelstorage, stage_host, writeback,
                     writeback_rows)


class SweepOperand:
    """src/executor.jl:3-11."""

    __slots__ = ("pm", "read", "write")

    def __init__(self, pm, read, write):
        self.pm = pm
        self.read = read
        self.write = write


def readpanels(pm):
    return SweepOperand(pm, True, False)


def writepanels(pm):
    return SweepOperand(pm, False, True)


def updatepanels(pm):
    return SweepOperand(pm, True, True)


class PanelSteps:
    """A step is one column panel; a panel may repeat, with a label per step
    (src/executor.jl:18-27)."""

    def __init__(self, panels, labels=None):
        if isinstance(panels, int):
            count = panels
            self.panels = list(range(count))
            self.labels = list(range(count))
        else:
            self.panels = [int(p) for p in panels]
            self.labels = list(labels) if labels is not None else list(self.panels)
        if len(self.labels) != len(self.panels):
            raise ArgumentError("a PanelSteps traversal needs one label per step")

    def __repr__(self):
        return f"PanelSteps({len(self.panels)} steps)"


class RowSteps:
    """A step is a block of rows spanning every panel (src/executor.jl:29-36).
    Blocks are (start, stop) pairs, 0-based half-open."""

    def __init__(self, N, height=None, blocks=None):
        if blocks is not None:
            self.blocks = [tuple(b) for b in blocks]
            return
        N = int(N)
        h = min(max(int(height), 1), N)
        self.blocks = [(top, min(top + h, N)) for top in range(0, N, h)]

    def __repr__(self):
        return f"RowSteps({len(self.blocks)} blocks)"


def rowtraversal(pm):
    """src/executor.jl:38."""
    return RowSteps(pm.N, _plan.row_block_height(pm.N, pm.k, pm.w))


def nsteps(t):
    return len(t.panels) if isinstance(t, PanelSteps) else len(t.blocks)


def steplabel(t, step):
    return t.labels[step] if isinstance(t, PanelSteps) else step


def repeats_a_panel(t):
    return isinstance(t, PanelSteps) and len(set(t.panels)) != len(t.panels)


def buffer_dims(t, pm):
    """src/executor.jl:49-50."""
    if isinstance(t, PanelSteps):
        return (pm.N, pm.w)
    b = t.blocks[0]
    return (b[1] - b[0], pm.k)


def assert_sweepable(t, X, Y):
    if isinstance(t, PanelSteps):
        assert_same_panels(X, Y)
    else:
        assert_same_rows(X, Y)


def step_view(t, buffer, step, pm):
    """The part of a staging buffer one step uses (src/executor.jl:55-56)."""
    if isinstance(t, PanelSteps):
        return buffer.view(None, (0, pm.panels[t.panels[step]].width))
    b = t.blocks[step]
    return buffer.view((0, b[1] - b[0]), None)


def resident_view(t, step, pm):
    """A step of a device-resident matrix: a view of the matrix itself."""
    if isinstance(t, PanelSteps):
        return panelstorage(pm.panels[t.panels[step]])
    b = t.blocks[step]
    return pm.device.view(b, None)


def stage_in(t, dst, pm, step, queue, via=None):
    if isinstance(t, PanelSteps):
        return materialize(dst, pm.panels[t.panels[step]], pm.plan, queue=queue, via=via)
    return materialize_rows(dst, pm, t.blocks[step], queue=queue, via=via)


def stage_out(t, pm, src, step, queue, via=None):
    if isinstance(t, PanelSteps):
        return writeback(pm.panels[t.panels[step]], src, pm.plan, queue=queue, via=via)
    return writeback_rows(pm, src, t.blocks[step], queue=queue, via=via)


def steppanels(t, pm, step):
    """Panels one step touches: one, or all of them for a row block."""
    if isinstance(t, PanelSteps):
        return (pm.panels[t.panels[step]],)
    return pm.panels


def stage_step(t, pm, step):
    for panel in steppanels(t, pm, step):
        stage_host(panel, pm.plan)


def release_step(t, operands, step, hold):
    """The pins a step took become the event its copies finish on."""
    for op in operands:
        for panel in steppanels(t, op.pm, step):
            _plan.unpin(op.pm.plan, panel, hold)


def clear_pins(operands):
    plan = operands[0].pm.plan
    with plan.hostpool.lock:
        for op in operands:
            for panel in op.pm.panels:
                panel.pins = 0


_FAILED = object()

# Observers of the sweeps started while trace_sweeps() is open: each gets one
# record per sweep, before its first step is issued.
_observers = []


@contextlib.contextmanager
def trace_sweeps():
    """Collects a record of every sweep issued inside the block, in order:
    the traversal ("rows", with each step's (start, stop) rows, or "panels",
    with each step's (panel, label)), each operand's mode ("read", "write" or
    "update") and whether it was staged, and the buffer count. The schedule
    of an operation is then comparable step for step with the reference's."""
    records = []
    _observers.append(records.append)
    try:
        yield records
    finally:
        _observers.remove(records.append)


def _trace(steps, operands, staged, nb):
    if isinstance(steps, PanelSteps):
        kind, seq = "panels", tuple(zip(steps.panels, steps.labels))
    else:
        kind, seq = "rows", tuple(steps.blocks)
    modes = tuple("update" if op.read and op.write else "read" if op.read else "write" for op in operands)
    record = {"traversal": kind, "steps": seq, "modes": modes, "staged": tuple(staged), "nbuffers": nb}
    for observe in list(_observers):
        observe(record)


class StageChannel:
    """The producer half of a sweep (src/executor.jl:102-111): stages every
    panel of each step, at most `depth` steps ahead of the issue loop."""

    def __init__(self, traversal, operands, depth=_plan.FEED_DEPTH):
        self._q = _queue.Queue(maxsize=max(int(depth), 1))
        self._closed = threading.Event()
        self._error = None
        self._traversal = traversal
        self._operands = operands
        self._thread = threading.Thread(target=self._run, name="funicular-producer", daemon=True)
        self._thread.start()

    def _put(self, item):
        while True:
            try:
                self._q.put(item, timeout=0.05)
                return True
            except _queue.Full:
                if self._closed.is_set():
                    return False

    def _run(self):
        try:
            for step in range(nsteps(self._traversal)):
                if self._closed.is_set():
                    return
                for op in self._operands:
                    stage_step(self._traversal, op.pm, step)
                if not self._put(step):
                    return
        except BaseException as err:  # surfaces on the issuing thread
            self._error = err
            self._put(_FAILED)

    def take(self):
        while True:
            try:
                item = self._q.get(timeout=0.05)
                break
            except _queue.Empty:
                if not self._thread.is_alive() and self._q.empty():
                    if self._error is not None:
                        raise self._error
                    raise RuntimeError("the sweep's producer stopped without staging every step")
        if item is _FAILED:
            raise self._error
        return item

    def available(self):
        return self._q.qsize()

    def close(self):
        self._closed.set()
        self._thread.join()


def stage_channel(traversal, operands, depth=_plan.FEED_DEPTH):
    return StageChannel(traversal, operands, depth)


def sync_quietly(backend, queues):
    for q in queues:
        try:
            backend.sync_queue(q)
        except Exception:
            pass


def sweep(f, *operands, nbuffers=None, traversal=None):
    """Calls f(label, *buffers) once per step with one device view per operand
    (src/executor.jl:165-239)."""
    if not operands:
        raise ArgumentError("sweep was given no operand; pass each panel matrix as readpanels(pm), "
                            "writepanels(pm) or updatepanels(pm)")
    reference = operands[0].pm
    plan = reference.plan
    steps = PanelSteps(npanels(reference)) if traversal is None else traversal
    for op in operands:
        assert_sweepable(steps, reference, op.pm)
        if op.pm.plan is not plan:
            raise ArgumentError("the operands of a sweep come from different ResidencyPlans; a sweep stages "
                                "through one plan's buffers, so build its matrices from a single plan")
        if op.write:
            assert_writable(op.pm)
    for i in range(1, len(operands)):
        for l in range(i):
            if operands[i].pm is operands[l].pm:
                raise ArgumentError("a panel matrix appears twice among the operands of one sweep; to read and "
                                    "write it in the same sweep, pass it once, as updatepanels(pm)")
    requested = plan.nbuffers if nbuffers is None else int(nbuffers)
    if requested < 1:
        raise ArgumentError(f"a sweep needs one staging buffer or more per operand; nbuffers = {requested}")
    writing = any(op.write for op in operands)
    if writing and repeats_a_panel(steps):
        raise ArgumentError("the traversal returns to a panel it has visited and the sweep writes back: the later "
                            "write-back would replace the earlier one. A traversal may repeat panels only when "
                            "every operand is read-only")

    backend = plan.backend
    count = nsteps(steps)
    if count == 0:
        return None
    nb = min(requested, count)
    staged = [not op.pm.resident for op in operands]
    if _observers:
        _trace(steps, operands, staged, nb)
    dims = [buffer_dims(steps, op.pm) for op in operands]
    buffers = []
    staging = []
    try:
        for op, d, s in zip(operands, dims, staged):
            buffers.append(_plan.checkout_device_buffers(plan, op.pm.dtype, d, nb) if s else None)
            staging.append(checkout_staging_buffers(plan, op.pm, d, nb) if s else None)
    except BaseException:
        for b in buffers:
            if b is not None:
                _plan.checkin_device_buffers(plan, b)
        for b in staging:
            if b is not None:
                _plan.checkin_device_buffers(plan, b)
        raise
    up, compute, down = backend.make_sweep_queues()
    slotfree = [None] * nb
    feed = stage_channel(steps, operands)
    failed = False
    try:
        for _ in range(count):
            step = feed.take()
            slot = step % nb
            targets = []
            vias = []
            for op, b, st, s in zip(operands, buffers, staging, staged):
                if s:
                    targets.append(step_view(steps, b[slot], step, op.pm))
                    vias.append(None if st is None else step_view(steps, st[slot], step, op.pm))
                else:
                    targets.append(resident_view(steps, step, op.pm))
                    vias.append(None)
            label = steplabel(steps, step)

            free = slotfree[slot]
            if free is not None:
                backend.wait_event(up, free)
            for op, target, via, s in zip(operands, targets, vias, staged):
                if op.read and s:
                    stage_in(steps, target, op.pm, step, up, via=via)
            backend.wait_event(compute, backend.record_event(up))
            done = backend.submit(compute, lambda: f(label, *targets))
            if writing:
                backend.wait_event(down, done)
                for op, target, via, s in zip(operands, targets, vias, staged):
                    if op.write:
                        if s:
                            stage_out(steps, op.pm, target, step, down, via=via)
                        else:
                            _mark_written(steps, op.pm, step)
                done = backend.record_event(down)
            slotfree[slot] = done
            release_step(steps, operands, step, done)
        # Compute first so the caller sees a failure of the panel function.
        backend.sync_queue(compute)
        backend.sync_queue(up)
        backend.sync_queue(down)
    except BaseException:
        failed = True
        raise
    finally:
        feed.close()
        sync_quietly(backend, (up, compute, down))
        if failed:
            # Once the producer has stopped, so that it pins nothing more.
            clear_pins(operands)
        for b in buffers:
            if b is not None:
                _plan.checkin_device_buffers(plan, b)
        for b in staging:
            if b is not None:
                _plan.checkin_device_buffers(plan, b)
    return None


def _mark_written(t, pm, step):
    for panel in steppanels(t, pm, step):
        panel.dirty = True
        panel.epoch += 1


def foreachpanel(f, pm, write=True, nbuffers=None):
    """src/executor.jl:241-267: f(j, panel) on every panel in order."""
    return sweep(f, updatepanels(pm) if write else readpanels(pm), nbuffers=nbuffers)
