"""GPU parity of the batched form: every job of a batch (one kernel sequence
for all screens) equals the plain pipe on that screen alone and the oracle,
bit for bit, in every raster mode."""
import sys

import numpy as np
import pytest

import piko_amd
from piko_amd import scenes

pytestmark = pytest.mark.gpu

# auto: a batch's default (the tiled path); queue: its default coverage
# (spans for compact triangles) and parts (whole tiles), queue-walk: every
# int32-path triangle walked, queue-parts: tiles cut into parts of 64 entries
# that meet in the merge launch
MODES = {"binned": piko_amd.RASTER_BINNED, "direct": piko_amd.RASTER_DIRECT, "tiled": piko_amd.RASTER_TILED,
         "queue": piko_amd.RASTER_QUEUE, "queue-walk": piko_amd.RASTER_QUEUE, "queue-parts": piko_amd.RASTER_QUEUE,
         "auto": piko_amd.RASTER_AUTO}


def cover(mode):
    return piko_amd.COVER_WALK if mode == "queue-walk" else piko_amd.COVER_AUTO


def part_size(mode):
    return 64 if mode == "queue-parts" else 0


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain(scene, T, S=8):
    with piko_amd.RasterPipe().allocate(scene, T, S) as pipe:
        pipe.prepare()
        pipe.run_single()
        return pipe.pikoScreen.getData(), pipe.get_depth()


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_of_uneven_scenes_matches_plain_and_oracle(oracle_mod, mode):
    rng = np.random.default_rng(5)
    scs, Ts = [], []
    for k in range(14):
        W = int(rng.integers(40, 420))
        H = int(rng.integers(30, 300))
        n = int(rng.integers(0, 6000)) if k != 3 else 0
        scs.append(scenes.tris_uniform(2000 + k, n, W, H))
        Ts.append([8, 16, 32, 64][k % 4])
    pos, attr, parts = piko_amd.concat_scenes(scs)
    log("batch of %d scenes, %d triangles, mode %s" % (len(scs), pos.shape[0], mode))
    with piko_amd.BatchPipe(pos, attr, "pretransformed", max_jobs=64, max_tris=1000, max_pixels=4096,
                            raster_mode=MODES[mode], queue_cover=cover(mode), part_pairs=part_size(mode)) as bp:
        bp.submit([(n, base, T, piko_amd.ConstantState.for_scene(s)) for (n, base, s), T in zip(parts, Ts)])
        bp.wait()
        for j, (s, T) in enumerate(zip(scs, Ts)):
            c, d = bp.color(j), bp.depth(j)
            pc, pd = plain(s, T)
            assert np.array_equal(c, pc) and np.array_equal(d.view(np.int32), pd.view(np.int32)), j
            oc, od, _ = oracle_mod.frame(s, 8, T)
            assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32)), j
        # the same batch re-run gives the same frames
        c0 = bp.color(0).copy()
        bp.run_frames(3)
        assert np.array_equal(bp.color(0), c0)


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_world_jobs_share_resident_geometry(oracle_mod, mode):
    base = scenes.heightfield(seed=4, grid=120, W=320, H=180)
    cams = [((0.0, 0.3, 1.0), (0.0, -0.3, 0.0), 320, 180, 16),
            ((0.4, 0.5, 1.2), (0.0, 0.0, 0.0), 480, 270, 8),
            ((-0.6, 0.8, 0.9), (0.1, -0.1, 0.0), 200, 200, 32)]
    log("world batch, mode %s" % mode)
    with piko_amd.BatchPipe(base.pos, base.nrm, "world", max_jobs=8, max_tris=base.n, max_pixels=1 << 16,
                            raster_mode=MODES[mode], queue_cover=cover(mode), part_pairs=part_size(mode)) as bp:
        jobs, views = [], []
        for eye, tgt, W, H, T in cams:
            v = scenes.WorldScene(base.pos, base.nrm, W, H,
                                  scenes.view_proj(eye, tgt, scenes.HF_UP, 60.0, W / H, 0.02, 4.0),
                                  base.light, base.matcol, base.ambient)
            views.append((v, T))
            jobs.append((base.n, 0, T, piko_amd.ConstantState.for_scene(v)))
        bp.submit(jobs)
        bp.wait()
        for j, (v, T) in enumerate(views):
            oc, od, _ = oracle_mod.frame(v, 8, T)
            assert np.array_equal(bp.color(j), oc) and np.array_equal(bp.depth(j).view(np.int32), od.view(np.int32))


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_resubmit_different_sizes(mode):
    a = [scenes.tris_uniform(30 + k, 2000 * (k + 1), 100 + 50 * k, 80 + 30 * k) for k in range(3)]
    pos, attr, parts = piko_amd.concat_scenes(a)
    log("resubmit, mode %s" % mode)
    with piko_amd.BatchPipe(pos, attr, "pretransformed", max_jobs=8, max_tris=10, max_pixels=100,
                            raster_mode=MODES[mode], queue_cover=cover(mode), part_pairs=part_size(mode)) as bp:
        for order in ([0, 1, 2], [2, 0], [1]):
            bp.submit([(parts[k][0], parts[k][1], 16, piko_amd.ConstantState.for_scene(a[k])) for k in order])
            for j, k in enumerate(order):
                pc, _ = plain(a[k], 16)
                assert np.array_equal(bp.color(j), pc)
            # several frames of the job set in one call
            bp.run_frames(3)
            for j, k in enumerate(order):
                pc, _ = plain(a[k], 16)
                assert np.array_equal(bp.color(j), pc)
