"""GPU serving: requests replayed with continuous batching give, for every
request, exactly the frame it gives alone and the oracle's frame."""
import ctypes
import sys

import numpy as np
import pytest

import piko_amd
from piko_amd import scenes, serve

pytestmark = pytest.mark.gpu

SMALL_RES = ((160, 90), (200, 120), (96, 96))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def view_scene(pos, nrm, base, count, sc, view):
    b, n = int(base[sc]), int(count[sc])
    return scenes.WorldScene(pos[b:b + n], nrm[b:b + n], view.W, view.H, view.view_proj, view.light, view.matcol,
                             view.ambient)


def test_same_alone_and_in_the_mix(oracle_mod):
    pos, nrm, base, count = serve.resident_scenes(6, 3000, seed=3)
    reqs, meta = serve.poisson_trace(30, 50000.0, 6, seed=5, resolutions=SMALL_RES)
    with serve.Server(pos, nrm, base, count, pipes=2, max_jobs=8, budget_bytes=64 << 20) as s:
        done, st, cols = s.replay(reqs, keep=True)
        log("replay: %s" % st)
        assert st["requests"] == 30 and st["batches"] < 30  # requests were batched together
        assert np.all(done >= np.array([r.arrival_s for r in reqs]))
        for i in range(len(reqs)):
            one = (serve.Request * 1)()
            one[0] = reqs[i]
            one[0].arrival_s = 0.0
            _, st1, c1 = s.replay(one, keep=True)
            assert st1["batches"] == 1
            assert np.array_equal(c1[0], cols[i]), i
    for i in range(4):
        sc, view = meta[i]
        oc, od, _ = oracle_mod.frame(view_scene(pos, nrm, base, count, sc, view), 8, 8)
        assert np.array_equal(cols[i], oc), i


def test_admission_refuses_what_never_fits():
    pos, nrm, base, count = serve.resident_scenes(2, 2000, seed=9)
    reqs, _ = serve.poisson_trace(3, 1000.0, 2, seed=1, resolutions=((1920, 1080),))
    with serve.Server(pos, nrm, base, count, pipes=1, max_jobs=4, budget_bytes=1 << 20) as s:
        with pytest.raises(piko_amd.PikoError) as e:
            s.replay(reqs)
        assert e.value.status == -6


def test_pipes_fall_back_when_up_front_sizing_does_not_fit(oracle_mod):
    # a request cap and batch budget far beyond device memory: the budget is
    # cut to each pipe's share of the free memory, the first pipe is sized to
    # it up front, and the second no longer fits beside it, so it starts at
    # one request and grows with its batches; the server still starts and
    # every frame is the oracle's
    pos, nrm, base, count = serve.resident_scenes(4, 3000, seed=4)
    reqs, meta = serve.poisson_trace(12, 50000.0, 4, seed=6, resolutions=SMALL_RES)
    with serve.Server(pos, nrm, base, count, pipes=2, max_jobs=1 << 16, budget_bytes=1 << 44) as s:
        done, st, cols = s.replay(reqs, keep=True)
        log("replay: %s" % st)
    for i in range(3):
        sc, view = meta[i]
        oc, od, _ = oracle_mod.frame(view_scene(pos, nrm, base, count, sc, view), 8, 8)
        assert np.array_equal(cols[i], oc), i


def test_online_submit_while_batches_run():
    # requests submitted from a client thread while earlier batches run:
    # every frame equals the same request replayed alone
    import threading
    pos, nrm, base, count = serve.resident_scenes(6, 3000, seed=8)
    reqs, _ = serve.poisson_trace(24, 50000.0, 6, seed=9, resolutions=SMALL_RES)
    with serve.Server(pos, nrm, base, count, pipes=2, max_jobs=4, budget_bytes=64 << 20) as s:
        got = {}
        def client():
            for i, r in enumerate(reqs):
                rid, col = s.submit(r)
                got[rid] = (i, col)
        th = threading.Thread(target=client)
        th.start()
        done = []
        while len(done) < len(reqs):
            done += s.poll(16, 50)
        th.join()
        assert all(st == 0 for _, _, st in done)
        for rid, _, _ in done:
            i, col = got[rid]
            one = (serve.Request * 1)()
            one[0] = reqs[i]
            one[0].arrival_s = 0.0
            _, _, c1 = s.replay(one, keep=True)
            assert np.array_equal(c1[0], col), i


def test_admission_bounded_by_device_memory():
    # a batch budget far beyond the device: admission uses each pipe's share
    # of the free device memory instead, so a request that could never be
    # rendered beside the other pipes (a 40000 x 40000 screen, 25.6 GB of
    # framebuffer and visibility words, against 16 pipes' shares of one GPU)
    # is refused at submission, while an ordinary one is served
    pos, nrm, base, count = serve.resident_scenes(2, 2000, seed=11)
    reqs, _ = serve.poisson_trace(2, 1000.0, 2, seed=3, resolutions=SMALL_RES)
    with serve.Server(pos, nrm, base, count, pipes=16, max_jobs=4, budget_bytes=1 << 60) as s:
        huge = serve.Request()
        ctypes.pointer(huge)[0] = reqs[0]
        huge.cs.screenSizeX = huge.cs.screenSizeY = 40000
        with pytest.raises(piko_amd.PikoError) as e:
            s.submit(huge, keep=False)
        assert e.value.status == -6
        rid, col = s.submit(reqs[1])
        done = []
        while not done:
            done = s.poll(4, 100)
        assert done[0][0] == rid and done[0][2] == 0
