"""Serving on the served configuration's own mix: 64 resident world-space
scenes (triangle soups and height fields) built scene by scene in device
memory, scene 0 at the full per-scene size of the ~40 GB configuration
(8.7M triangles) and the rest smaller; a Poisson trace at 2,000 req/s with
960x540, 1280x720 and 1920x1080 screens and cameras on a sphere of radius
2-5 around the scene.  Every request's frame equals the same request
rendered alone, both when replayed and when submitted online; 8 sampled
requests, scene 0's among them, equal the oracle."""
import sys
import time

import numpy as np
import pytest

import piko_amd
from piko_amd import scenes, serve

pytestmark = pytest.mark.gpu

FULL = 8_700_000  # ~40 GB / 64 scenes / 72 bytes per triangle


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config5_mix_alone_online_and_oracle(oracle_mod):
    sizes = [FULL] + [20_000] * 63
    ds = serve.DeviceScenes(64, sizes=sizes, log=log)
    try:
        reqs, meta = serve.poisson_trace(48, 2000.0, 64, seed=21)
        # requests 0, 7 and 20 render the full-size scene
        for i in (0, 7, 20):
            reqs[i].scene = 0
            meta[i] = (0, meta[i][1])
        assert {(r.cs.screenSizeX, r.cs.screenSizeY) for r in reqs} == set(serve.RESOLUTIONS)
        log("config 5: %.2f GB resident, %d requests" % (ds.nbytes / 1e9, len(reqs)))
        with serve.Server(ds, pipes=4, max_jobs=16, budget_bytes=6 << 30) as s:
            done, st, cols = s.replay(reqs, keep=True)
            log("replay: %s" % st)
            assert st["batches"] < len(reqs)  # requests shared batches
            for i in range(len(reqs)):
                one = (serve.Request * 1)()
                one[0] = reqs[i]
                one[0].arrival_s = 0.0
                _, st1, c1 = s.replay(one, keep=True)
                assert st1["batches"] == 1
                assert np.array_equal(c1[0], cols[i]), i
            log("every request equals itself alone")
            # online: all submitted at once from this thread, collected by polling
            out = {}
            for i, r in enumerate(reqs):
                rid, col = s.submit(r)
                out[rid] = (i, col)
            seen = 0
            t0 = time.time()
            while seen < len(reqs):
                for rid, _, status in s.poll(64, 100):
                    assert status == 0
                    i, col = out[rid]
                    assert np.array_equal(col, cols[i]), i
                    seen += 1
                assert time.time() - t0 < 120
            log("online submissions equal the replay")
        for i in (0, 7, 3, 11, 20, 25, 33, 47):
            sc, view = meta[i]
            pos, nrm = ds.host_scene(sc)
            world = scenes.WorldScene(pos, nrm, view.W, view.H, view.view_proj, view.light, view.matcol,
                                      view.ambient)
            oc, od, _ = oracle_mod.frame(world, 8, 8, threads=16)
            assert np.array_equal(cols[i], oc), (i, sc, int((cols[i] != oc).sum()))
            log("request %d (scene %d, %dx%d) equals the oracle" % (i, sc, view.W, view.H))
    finally:
        ds.free()
