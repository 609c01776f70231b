"""GPU parity: the HIP Reyes pipeline (split -> dice -> shade -> resolve)
against the oracle through the C-ABI, bit for bit: split leaves (control
points and the reference's nu/nv), every sample's depth, the resolved FP32
RGBA and the per-pixel depth.  Every frame is checked with both sample
compositing paths: global (a 64-bit atomicMin per covered sample in a
visibility buffer, unless the sample's key already wins) and tiled (per-tile LDS visibility, 16- and 32-pixel
tiles)."""
import sys

import numpy as np
import pytest

from piko_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sphere(nlat, nlon, W, H, **kw):
    return scenes.sphere_patches(seed=3, nlat=nlat, nlon=nlon, W=W, H=H, **kw)


SAMPLINGS = (("global", 0), ("tiled", 16), ("tiled", 32))


def gpu_reyes(scene, **alloc):
    import piko_amd
    with piko_amd.ReyesPipe().allocate(scene, keep_samples=True, **alloc) as pipe:
        pipe.prepare()
        pipe.run_single()
        return pipe.color(), pipe.depth(), pipe.samples(), pipe.leaves(), pipe.stats()


def assert_reyes_equal(scene, oracle_mod, check_leaves=True, samplings=SAMPLINGS, **alloc):
    import piko_amd
    orgba, oz, nup = oracle_mod.reyes_frame(scene)
    if check_leaves:
        ocp, onu, onv = oracle_mod.reyes_split(scene.cp, scene.W, scene.H, scene.view_proj, scene.threshold)
    for mode, tile in samplings:
        log("case P=%d W=%d H=%d thr=%g dice=%d spp=%d disp=%g %s sampling=%s tile=%d" % (
            scene.n, scene.W, scene.H, scene.threshold, scene.dice, scene.spp, scene.disp, alloc, mode, tile))
        sampling = piko_amd.SAMPLE_TILED if mode == "tiled" else piko_amd.SAMPLE_GLOBAL
        rgba, depth, z, (lcp, lnu, lnv), st = gpu_reyes(scene, sampling=sampling, tile=tile, **alloc)
        if check_leaves:
            assert lcp.shape == ocp.shape, "leaf count %d vs oracle %d" % (lcp.shape[0], ocp.shape[0])
            assert np.array_equal(lcp.view(np.int32), ocp.view(np.int32)), "leaf control points differ"
            assert np.array_equal(lnu, onu) and np.array_equal(lnv, onv), "leaf nu/nv differ"
        assert st["micropolygons"] == nup, "micropolygons %d vs oracle %d" % (st["micropolygons"], nup)
        bad = np.argwhere(z != oz)
        assert bad.size == 0, "%s: sample depth differs at %d samples, first %s" % (mode, len(bad), bad[0])
        bad = np.argwhere(np.any(rgba.view(np.int32) != orgba.view(np.int32), axis=-1))
        assert bad.size == 0, "%s: colour differs at %d pixels, first %s: gpu %s oracle %s" % (
            mode, len(bad), bad[0], rgba[tuple(bad[0])], orgba[tuple(bad[0])])
        assert np.array_equal(depth.view(np.int32), oz.min(axis=2)), "%s: pixel depth differs" % mode
        assert np.isfinite(rgba).all()
        assert (st["tile_pairs"] > 0) == (mode == "tiled" and nup > 0)
    return rgba, st


@pytest.mark.parametrize("spp", [1, 4])
@pytest.mark.parametrize("disp", [0.0, 0.05])
def test_small_sphere(oracle_mod, spp, disp):
    assert_reyes_equal(sphere(8, 8, 320, 180, spp=spp, disp=disp), oracle_mod)


@pytest.mark.parametrize("dice", [1, 2, 4, 16, 32])
def test_dice_grids(oracle_mod, dice):
    assert_reyes_equal(sphere(6, 8, 256, 144, dice=dice, threshold=32.0, spp=1), oracle_mod)


def test_reference_dice_mode(oracle_mod):
    # the reference's own settings: 4-px split threshold, (nu+1)x(nv+1) grids
    # from the leaf extents, one sample at the pixel centre, no displacement
    assert_reyes_equal(sphere(4, 8, 160, 90, dice=0, threshold=4.0, spp=1, disp=0.0), oracle_mod)


def test_config3_full(oracle_mod):
    # 4,096 patches (64x64 lat-long sphere, radial noise +-0.05, seed 3),
    # split below 16 px, 8x8 dice, displacement, 4 samples, 1280x720 FP32
    sc = scenes.sphere_patches()
    rgba, st = assert_reyes_equal(sc, oracle_mod)
    log("config 3: leaves %d micropolygons %d levels %d" % (st["leaves"], st["micropolygons"], st["levels"]))
    assert st["levels"] >= 5


@pytest.mark.parametrize("rpw", [1, 3, 64])
def test_roots_per_wave_invariance(oracle_mod, rpw):
    # one root per wave: a root of this close-up view emits 163 leaves
    # against a first leaf region of 64, so the pipe must grow its regions
    # and re-run; the frame is the same for any grouping
    _, st = assert_reyes_equal(sphere(8, 16, 400, 300, eye=(0.0, 0.2, 2.2)), oracle_mod, roots_per_wave=rpw)
    if rpw == 1:
        assert st["capacity_reruns"] >= 1


def test_frames_in_one_call_regrow(oracle_mod):
    # several frames in one call run each frame's split beside the previous
    # frame's dice, alternating two slots of leaves and ranks; this close-up
    # view overflows the first leaf regions, so the slots are freed, grown
    # and the frames re-run
    import piko_amd
    sc = sphere(8, 16, 400, 300, eye=(0.0, 0.2, 2.2))
    orgba, oz, _ = oracle_mod.reyes_frame(sc)
    ocp, onu, onv = oracle_mod.reyes_split(sc.cp, sc.W, sc.H, sc.view_proj, sc.threshold)
    log("frames in one call with regrowth: P=%d" % sc.n)
    with piko_amd.ReyesPipe().allocate(sc, keep_samples=True, sampling=piko_amd.SAMPLE_GLOBAL,
                                       roots_per_wave=1) as pipe:
        pipe.prepare()
        pipe.run_frames(4)
        st = pipe.stats()
        assert st["capacity_reruns"] >= 1
        assert np.array_equal(pipe.color().view(np.int32), orgba.view(np.int32)), "colour differs"
        assert np.array_equal(pipe.samples(), oz), "sample depth differs"
        lcp, lnu, lnv = pipe.leaves()
        assert np.array_equal(lcp.view(np.int32), ocp.view(np.int32)), "leaves differ"
        assert np.array_equal(lnu, onu) and np.array_equal(lnv, onv)
        pipe.run_frames(3)
        assert np.array_equal(pipe.color().view(np.int32), orgba.view(np.int32)), "colour differs on later frames"


def test_more_split_waves_than_rank_registers(oracle_mod):
    # 4,608 patches, one per split wave: the rank kernel's threads each own a
    # run of 5 waves, past the 4 it keeps in registers, so the loop form runs
    sc = sphere(72, 64, 320, 240)
    _, st = assert_reyes_equal(sc, oracle_mod, check_leaves=False, samplings=(("global", 0),), roots_per_wave=1)
    log("4608 patches: leaves %d micropolygons %d" % (st["leaves"], st["micropolygons"]))
    assert sc.n == 4608 and st["leaves"] > 0


@pytest.mark.parametrize("mode", ["global", "tiled"])
def test_frames_repeat_and_camera_change(oracle_mod, mode):
    import piko_amd
    a = sphere(8, 8, 256, 256)
    b = sphere(8, 8, 256, 256, eye=(1.5, -0.5, 2.0))
    sampling = piko_amd.SAMPLE_TILED if mode == "tiled" else piko_amd.SAMPLE_GLOBAL
    log("repeat and camera change, sampling=%s" % mode)
    with piko_amd.ReyesPipe().allocate(a, keep_samples=True, sampling=sampling) as pipe:
        pipe.prepare()
        pipe.run_frames(3)
        c3 = pipe.color()
        oa, _, _ = oracle_mod.reyes_frame(a)
        assert np.array_equal(c3.view(np.int32), oa.view(np.int32)), "repeated frames differ"
        pipe.prepare(b)
        pipe.run_single()
        ob, oz, _ = oracle_mod.reyes_frame(b)
        assert np.array_equal(pipe.color().view(np.int32), ob.view(np.int32)), "new camera differs"
        assert np.array_equal(pipe.samples(), oz)
        # several frames in one call: on the global path each resolve runs on
        # a second stream beside the next frame's split
        pipe.run_frames(5)
        assert np.array_equal(pipe.color().view(np.int32), ob.view(np.int32)), "frames in one call differ"
        assert np.array_equal(pipe.samples(), oz)


def test_tiled_lists_grow(oracle_mod):
    # 8 patches seen from close by: the first (tile, leaf) list capacity
    # (4096 entries) is too small, so the lists grow and the frame is
    # re-run, with the same image
    sc = sphere(2, 4, 480, 270, eye=(0.0, 0.2, 1.6))
    _, st = assert_reyes_equal(sc, oracle_mod, samplings=(("tiled", 16),), roots_per_wave=8)
    log("pairs %d leaves %d reruns %d" % (st["tile_pairs"], st["leaves"], st["capacity_reruns"]))
    assert st["tile_pairs"] > 4096 and st["capacity_reruns"] >= 1


def test_tiled_tile_grows_with_screen(oracle_mod):
    # 2560x1440 in 16-pixel tiles would be 14,400 tiles, more than the 8,192
    # the binning histograms hold: the tiled path uses 32-pixel tiles there
    sc = sphere(4, 8, 2560, 1440, spp=1)
    assert_reyes_equal(sc, oracle_mod, check_leaves=False, samplings=(("tiled", 16),))


@pytest.mark.parametrize("mode", ["global", "tiled"])
def test_no_patches(mode):
    # an empty patch list: every pixel is the clear colour (fromABGR of
    # 0xff663313) at depth 1.0, on both sampling paths
    import piko_amd
    sc = sphere(2, 2, 96, 64)
    sc.cp = sc.cp[:0]
    sampling = piko_amd.SAMPLE_TILED if mode == "tiled" else piko_amd.SAMPLE_GLOBAL
    with piko_amd.ReyesPipe().allocate(sc, sampling=sampling) as pipe:
        pipe.prepare()
        pipe.run_frames(2)
        rgba, depth, st = pipe.color(), pipe.depth(), pipe.stats()
    clear = np.array([0x13 / 255.0, 0x33 / 255.0, 0x66 / 255.0, 1.0], np.float32)
    assert st["leaves"] == 0 and st["micropolygons"] == 0
    assert np.array_equal(rgba, np.broadcast_to(clear, rgba.shape)) and np.all(depth == 1.0)


def test_frames_without_kept_samples(oracle_mod):
    # keep_samples off (the benchmark's setting): the resolved frame is the
    # oracle's on both sampling paths
    import piko_amd
    sc = sphere(8, 8, 320, 180)
    orgba, oz, _ = oracle_mod.reyes_frame(sc)
    for sampling in (piko_amd.SAMPLE_GLOBAL, piko_amd.SAMPLE_TILED):
        with piko_amd.ReyesPipe().allocate(sc, sampling=sampling) as pipe:
            pipe.prepare()
            pipe.run_frames(3)
            assert np.array_equal(pipe.color().view(np.int32), orgba.view(np.int32)), sampling
            assert np.array_equal(pipe.depth().view(np.int32), oz.min(axis=2)), sampling


def test_invalid_options():
    import piko_amd
    sc = sphere(2, 2, 64, 64)
    for bad in (dict(dice=3), dict(spp=2), dict(threshold=0.0)):
        s2 = sphere(2, 2, 64, 64, **bad)
        with pytest.raises(piko_amd.PikoError):
            piko_amd.ReyesPipe().allocate(s2)
    with pytest.raises(piko_amd.PikoError):
        piko_amd.ReyesPipe().allocate(sc, max_splits=21)
    for bad in (dict(sampling=3), dict(sampling=-1), dict(tile=8), dict(tile=64)):
        with pytest.raises(piko_amd.PikoError):
            piko_amd.ReyesPipe().allocate(sc, **bad)
    # more than 8,192 tiles even at 32 pixels: refused for tiled sampling
    big = sphere(2, 2, 4096, 4096)
    with pytest.raises(piko_amd.PikoError):
        piko_amd.ReyesPipe().allocate(big, sampling=piko_amd.SAMPLE_TILED)
