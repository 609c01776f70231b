"""GPU parity on the batched configuration's own workload: 256 uniform-soup
scenes from seeds 1000-1255 (triangle counts log-uniform in [1K, 500K],
pixel counts log-uniform in [256^2, 2048^2], aspect 1, 4:3 or 16:9), all of
them in one batch with autotuned tile sizes, on the tiled path and on the
queue path (one persistent raster launch) in both of its coverage modes.  Every scene's colour words and
depth bits equal the oracle's (threaded, on the reference CPU schedule) and
its plain launch's."""
import sys

import numpy as np
import pytest

import piko_amd
from piko_amd import autotune, scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config4_every_scene_matches_oracle_and_plain(oracle_mod):
    spec = scenes.batch_spec()
    assert len(spec) == 256 and spec[0][0] == 1000 and spec[-1][0] == 1255
    scs = [scenes.tris_uniform(seed, n, W, H) for seed, n, W, H in spec]
    pos, attr, parts = piko_amd.concat_scenes(scs)
    ntris, npix = sum(s.n for s in scs), sum(s.W * s.H for s in scs)
    log("config 4: %d scenes, %d triangles, %d pixels; tuning tile sizes" % (len(scs), ntris, npix))
    tiles, _ = autotune.tune_batch(scs, pos, attr, parts, frames=1)
    jobs = [(n, base, T, piko_amd.ConstantState.for_scene(s)) for (n, base, s), T in zip(parts, tiles)]
    pipes = {}
    try:
        # the queue path with its defaults (whole tiles), and walked with
        # 2,048-entry parts (split tiles meet in the merge launch)
        for name, mode, qc, pp in (("tiled", piko_amd.RASTER_TILED, 0, 0), ("queue", piko_amd.RASTER_QUEUE, 0, 0),
                                   ("queue-walk", piko_amd.RASTER_QUEUE, piko_amd.COVER_WALK, 2048)):
            bp = piko_amd.BatchPipe(pos, attr, "pretransformed", max_jobs=len(scs), max_tris=ntris,
                                    max_pixels=npix, raster_mode=mode, queue_cover=qc, part_pairs=pp)
            pipes[name] = bp
            bp.submit(jobs)
            bp.wait()
            log("batch on the %s path done" % name)
        for j, (s, T) in enumerate(zip(scs, tiles)):
            oc, od, _ = oracle_mod.frame(s, 8, T, threads=16)
            for name, bp in pipes.items():
                c, d = bp.color(j), bp.depth(j)
                assert np.array_equal(c, oc), (name, j, int((c != oc).sum()))
                assert np.array_equal(d.view(np.int32), od.view(np.int32)), (name, j)
            with piko_amd.RasterPipe().allocate(s, T, 8, raster_mode=piko_amd.RASTER_DIRECT) as pipe:
                pipe.run_single()
                assert np.array_equal(pipe.pikoScreen.getData(), oc), ("plain", j)
            if j % 32 == 31:
                log("scenes 0..%d equal" % j)
    finally:
        for bp in pipes.values():
            bp.destroy()
