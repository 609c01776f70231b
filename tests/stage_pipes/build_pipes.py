"""Builds the test pipes under tests/stage_pipes (compiled stage code, as
the samples are built): edge/edge_pipe and edge/edge_pipe_bad_bin.

  python tests/stage_pipes/build_pipes.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "piko-public_amd"))
sys.path.insert(0, os.path.join(REPO, "piko-public_amd", "samples"))

PIPES = [
    ("edge", "edgePipe.h", "edge_pipe", {}),
    ("edge", "edgePipe.h", "edge_pipe_bad_bin", {"EDGE_BAD_BIN": "1"}),
]


def build_all(log=None):
    import build_samples
    return [build_samples.build_one(*p, log=log, root=HERE) for p in PIPES]


if __name__ == "__main__":
    for path in build_all(log=print):
        print(path)
