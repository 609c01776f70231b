"""GPU: compiled stage code on the edge cases (tests/stage_pipes/edge).

The edge pipe's raster stage fills boxes within its bins with a depth test
that lets a later box win a tie; its host program recomputes every pixel's
winner and depth in order and counts the pixels the GPU frame gets wrong.
The bad-bin build assigns the boxes from id 7 on to a bin past the grid
as well: the schedule drops those entries, the frame returns PIKO_ERR_INVALID and the program stops
with the message (no kernel faults: the entry is range-checked before it is
written)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PIPES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stage_pipes")


def run_pipe(binary, count):
    path = os.path.join(PIPES, binary)
    if not os.path.exists(path):
        pytest.fail("test pipe %s is not built (run __graft_entry__.build())" % binary)
    print("pipe", binary, count, flush=True)
    return subprocess.run([path, str(count)], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("count", [0, 1, 7, 3000])
def test_edge_pipe_frames(count):
    r = run_pipe("edge/edge_pipe", count)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    assert got["count"] == count and got["wrong"] == 0, got
    assert (got["covered"] > 0) == (count > 0)


def test_bin_outside_the_grid_is_reported():
    r = run_pipe("edge/edge_pipe_bad_bin", 100)
    assert r.returncode != 0
    assert "assignToBin outside the bin grid" in r.stderr and "invalid argument" in r.stderr, r.stderr
    # with boxes 0-6 only the same build runs clean
    r = run_pipe("edge/edge_pipe_bad_bin", 7)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1])["wrong"] == 0
