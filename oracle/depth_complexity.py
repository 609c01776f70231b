"""Depth complexity of the 2M-triangle height field at 1080p over camera
poses and noise lattice frequencies -- TEST INFRASTRUCTURE (the oracle's
vertex stage and exact coverage counts; nothing here runs on the GPU).

Depth complexity = fragments (covered pixel centres, before the depth test,
of every triangle the vertex stage emits) / covered pixels.  Prints a
markdown table.  Usage:

    python oracle/depth_complexity.py [freq,freq,...]
"""
import itertools
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]  # `oracle` is the package here
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "piko-public_amd"))

from oracle import oracle  # noqa: E402
from piko_amd import scenes  # noqa: E402

W, H, S = 1920, 1080, 8


def stats(hf, eye, target):
    vp = scenes.view_proj(eye, target, scenes.HF_UP, 60.0, W / H, 0.02, 4.0)
    sc = scenes.WorldScene(hf.pos, hf.nrm, W, H, vp, hf.light, hf.matcol, hf.ambient)
    st = oracle.vertex_stage(sc, S)
    frags, cov = oracle.fragment_stats(st, W, H, S)
    return frags, cov, int(st.keep.sum())


def main():
    freqs = [float(f) for f in sys.argv[1].split(",")] if len(sys.argv) > 1 else [4.0, 8.0, 16.0]
    print("| lattice cells per unit | eye | target | fragments | covered px | depth complexity | coverage | emitted |")
    print("|---|---|---|---|---|---|---|---|")
    for f in freqs:
        hf = scenes.heightfield(base_freq=f)
        for ey, ez, ty in itertools.product([0.1, 0.12, 0.2, 0.3], [0.7, 1.0], [-0.3]):
            eye, target = (0.0, ey, ez), (0.0, ty, 0.0)
            frags, cov, kept = stats(hf, eye, target)
            print("| %g | %s | %s | %d | %d | %.2f | %.3f | %d |" % (f, eye, target, frags, cov, frags / max(cov, 1),
                                                                   cov / (W * H), kept), flush=True)


if __name__ == "__main__":
    main()
