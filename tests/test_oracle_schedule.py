"""The oracle's schedule (oracle/schedule.py) pinned to what the reference
states about its own operations, so that the GPU schedule tests compare the
backend with the reference and not with a restatement of itself."""
import itertools

import pytest

from oracle import plan as P
from oracle import schedule as OS

SHAPES = [(1 << 20, 64, 8), (100003, 23, 7), (12, 20, 7), (1000, 1, 1), (33386248, 1024, 32)]


def _covers_rows_once(record, N):
    blocks = record["steps"]
    return blocks[0][0] == 0 and blocks[-1][1] == N and all(a[1] == b[0] for a, b in zip(blocks, blocks[1:]))


@pytest.mark.parametrize("N,k,w", SHAPES)
def test_cholqr2_is_four_sweeps_of_the_whole_matrix(N, k, w):
    # src/panelblas.jl:111 counts four full sweeps for the two passes; the
    # shift adds a third pass (:114-120, :135-138).
    for shifted, passes in ((False, 2), (True, 3)):
        ev = OS.cholqr2(N, k, w, shifted=shifted)
        sw = OS.sweeps(ev)
        assert len(sw) == 2 * passes
        assert all(r["traversal"] == "rows" and _covers_rows_once(r, N) for r in sw)
        # A pass reads Y for its Gram and updates it in its solve.
        assert [r["modes"] for r in sw] == [("read",), ("update",)] * passes
        t = OS.transfers(ev, N, k, w, 16)
        # Y moves up four times and down twice; per pass the Gram comes down
        # (collect_small, :61-65) and R goes up (:161).
        assert t["h2d_bytes"] == passes * (2 * N * k + k * k) * 16
        assert t["d2h_bytes"] == passes * (N * k + k * k) * 16


@pytest.mark.parametrize("N,k,w", SHAPES)
def test_gram_moves_each_operand_once(N, k, w):
    # src/panelblas.jl:14-20: gram(X) moves the whole matrix exactly once,
    # gram(X, Y) costs two uploads; the k x k result comes back once (:14).
    one = OS.transfers(OS.gram(N, k, w), N, k, w, 16)
    two = OS.transfers(OS.gram(N, k, w, two_operands=True), N, k, w, 16)
    assert one["h2d_bytes"] == N * k * 16 and two["h2d_bytes"] == 2 * N * k * 16
    assert one["d2h_bytes"] == two["d2h_bytes"] == k * k * 16
    # One 2-D copy per panel per row block (src/panels.jl:437-445), and the
    # accumulator's one copy back.
    p = len(P.panel_ranges(k, w))
    nblocks = len(P.row_blocks(N, P.row_block_height(N, k, w)))
    assert one["copies"] == p * nblocks + 1
    assert two["copies"] == 2 * p * nblocks + 1


@pytest.mark.parametrize("N,k,w", SHAPES[:4])
def test_project_uploads_p_times_p_plus_one_panels(N, k, w):
    # src/panelblas.jl:256-258: project uploads p (p + 1) panels and applies G
    # only p times; every block of S is assigned by exactly one step (:285).
    p = len(P.panel_ranges(k, w))
    ev = OS.project(N, k, w)
    (sw,) = OS.sweeps(ev)
    steps = sw["steps"]
    assert len(steps) == p * (p + 1)
    applying = [label for _, label in steps if label[0]]
    assert applying == [(True, j, j) for j in range(p)]
    blocks = sorted((label[1], label[2]) for _, label in steps if not label[0])
    assert blocks == sorted(itertools.product(range(p), range(p)))
    # A reducing step streams panel i against the image of panel j.
    assert all(panel == label[1] for panel, label in steps)
    t = OS.transfers(ev, N, k, w, 16)
    assert t["h2d_bytes"] == (p + 1) * N * k * 16  # panel j, then all of Q, for each j
    assert t["copies"] == p * (p + 1) + 1


@pytest.mark.parametrize("N,k,w", SHAPES[:4])
def test_panelmul_streams_x_up_and_y_down(N, k, w):
    # src/panelblas.jl:72-74: panels of X stream up, panels of Y back down.
    t = OS.transfers(OS.panelmul(N, k, w), N, k, w, 16)
    p = len(P.panel_ranges(k, w))
    assert t == {"h2d_bytes": N * k * 16, "d2h_bytes": N * k * 16, "copies": 2 * p}


def test_a_dropped_copy_or_sweep_is_caught():
    # The checks above fail for the plausible slips: a pass without its
    # solve's download, or a project without the image step.
    N, k, w = 1 << 20, 64, 8
    ev = OS.cholqr2(N, k, w)
    slipped = [dict(e, modes=("read",)) if isinstance(e, dict) and e["modes"] == ("update",) else e for e in ev]
    assert OS.transfers(slipped, N, k, w, 16)["d2h_bytes"] != 2 * (N * k + k * k) * 16
    (sw,) = OS.sweeps(OS.project(N, k, w))
    assert len([s for s in sw["steps"] if not s[1][0]]) != len(sw["steps"])
