"""Schedule arithmetic of the reference, restated for the oracle.

TEST INFRASTRUCTURE (see oracle/__init__.py). Integer arithmetic only; every
function cites the reference line it restates.
"""

COMPLEX128_SIZE = 16


def fld(a, b):
    return a // b


def cld(a, b):
    return -((-a) // b)


def choose_panel_width(N, k, itemsize, device_budget, nbuffers=2, extra_buffers=0, workspace_bytes=0,
                       target_panel_bytes=2 * 1024 ** 3):
    """src/plan.jl:112-123. Raises ValueError where the reference throws."""
    if N <= 0 or k <= 0:
        raise ValueError("N and k must be positive")
    buffers = nbuffers + extra_buffers
    if buffers < 1:
        raise ValueError("need at least one buffer")
    column = N * itemsize
    available = int(device_budget) - int(workspace_bytes)
    per_buffer = fld(available, buffers)
    if per_buffer < column:
        raise ValueError("budget below one column per buffer")
    w = min(k, max(1, fld(min(int(target_panel_bytes), per_buffer), column)))
    return cld(k, cld(k, w))


def row_block_height(N, k, w):
    """src/plan.jl:133: clamp(fld(N w, k), 1, N)."""
    return min(max(fld(N * w, k), 1), N)


def row_blocks(N, height):
    """src/executor.jl:31-36: RowSteps(N, height) as 0-based (start, stop)."""
    h = min(max(height, 1), N)
    return [(top, min(top + h, N)) for top in range(0, N, h)]


def panel_ranges(k, w):
    """src/panels.jl:70-71: columns of each panel, 0-based (start, stop)."""
    return [(c, min(c + w, k)) for c in range(0, k, w)]


def project_labels(p):
    """src/panelblas.jl:272-281 with 0-based panels."""
    panels, labels = [], []
    for j in range(p):
        panels.append(j)
        labels.append((True, j, j))
        for i in range(p):
            panels.append(i)
            labels.append((False, i, j))
    return panels, labels


def slot(step, nb):
    """src/executor.jl:199: mod1(step, nb) for 1-based steps, here 0-based."""
    return step % nb


def device_bytes_required(N, w, itemsize, nbuffers=2, extra_buffers=0, workspace_bytes=0):
    """src/plan.jl:99-101."""
    return (nbuffers + extra_buffers) * N * w * itemsize + workspace_bytes
