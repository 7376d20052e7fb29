"""The reference's step and transfer schedule for the panel BLAS, restated.

TEST INFRASTRUCTURE (see oracle/__init__.py). Given a matrix's shape, this
lists the sweeps an operation runs on the reference's host-resident path, in
order, each as the traversal it walks and the modes of its operands, and the
small k x k copies between them; `transfers` then counts the bytes and copy
calls the reference's device contract issues for them. Integer bookkeeping
only, each function citing the reference lines it follows.

A sweep record has the shape executor.trace_sweeps() produces:
  {"traversal": "rows",   "steps": ((start, stop), ...),   "modes": (...)}
  {"traversal": "panels", "steps": ((panel, label), ...), "modes": (...)}
and a small copy is ("h2d" | "d2h", nbytes_over_itemsize).
"""
from . import plan as P


def _modes(record):
    return record["modes"]


def row_sweep(N, k, w, *modes):
    """A sweep over the row traversal (src/executor.jl:29-38; height
    src/plan.jl:133)."""
    return {"traversal": "rows", "steps": tuple(P.row_blocks(N, P.row_block_height(N, k, w))),
            "modes": tuple(modes)}


def panel_sweep(panels, labels, *modes):
    """A sweep over PanelSteps(panels, labels) (src/executor.jl:18-27)."""
    return {"traversal": "panels", "steps": tuple(zip(panels, labels)), "modes": tuple(modes)}


def gram(N, k, w, two_operands=False):
    """gram(X) or gram(X, Y): one read row sweep over every operand, then the
    k x k accumulator copied back once (src/panelblas.jl:33-53, collect_small
    :61-65)."""
    modes = ("read", "read") if two_operands else ("read",)
    return [row_sweep(N, k, w, *modes), ("d2h", k * k)]


def rdiv_rows(N, k, w):
    """rdiv_rows!: R uploaded, then one update row sweep (src/panelblas.jl:
    157-169)."""
    return [("h2d", k * k), row_sweep(N, k, w, "update")]


def cholqr_pass(N, k, w):
    """src/panelblas.jl:141-146: gram, the host Cholesky, rdiv_rows!."""
    return gram(N, k, w) + rdiv_rows(N, k, w)


def cholqr2(N, k, w, shifted=False):
    """src/panelblas.jl:130-139: two passes, and a third when either
    Cholesky needed a shift."""
    events = cholqr_pass(N, k, w) + cholqr_pass(N, k, w)
    if shifted:
        events += cholqr_pass(N, k, w)
    return events


def project(N, k, w):
    """src/panelblas.jl:268-298: one read sweep of p (p + 1) panel steps,
    labelled (applying, i, j) with 0-based panels, then S copied back."""
    p = len(P.panel_ranges(k, w))
    panels, labels = P.project_labels(p)
    return [panel_sweep(panels, labels, "read"), ("d2h", k * k)]


def panelmul(N, k, w):
    """src/panelblas.jl:91-100: one sweep over the column panels, X read and
    Y written."""
    p = len(P.panel_ranges(k, w))
    return [panel_sweep(range(p), range(p), "read", "write")]


def sweeps(events):
    """The sweep records of a schedule, without the small copies."""
    return [e for e in events if isinstance(e, dict)]


def transfers(events, N, k, w, itemsize):
    """Bytes up, bytes down and copy calls of a schedule on a host-resident
    matrix of N rows and k columns in panels of w: a row-block step moves its
    rows of every panel, one 2-D copy per panel (src/panels.jl:437-458); a
    panel step moves one whole panel in one copy (:373-398); a read operand
    goes up before the step and a written one comes down after it
    (src/executor.jl:205-216)."""
    ranges = P.panel_ranges(k, w)
    up = down = copies = 0
    for e in events:
        if not isinstance(e, dict):
            kind, n = e
            if kind == "h2d":
                up += n * itemsize
            else:
                down += n * itemsize
            copies += 1
            continue
        for step in e["steps"]:
            if e["traversal"] == "rows":
                r0, r1 = step
                nbytes, ncopies = (r1 - r0) * k * itemsize, len(ranges)
            else:
                c0, c1 = ranges[step[0]]
                nbytes, ncopies = N * (c1 - c0) * itemsize, 1
            for mode in e["modes"]:
                if mode in ("read", "update"):
                    up += nbytes
                    copies += ncopies
                if mode in ("write", "update"):
                    down += nbytes
                    copies += ncopies
    return {"h2d_bytes": up, "d2h_bytes": down, "copies": copies}
