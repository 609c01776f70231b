"""GPU parity: the HIP raster pipeline against the oracle, bit for bit
(colour words, depth bits and the per-bin queues), through the C-ABI.  Every
frame is checked in every raster mode: binned (queues sorted by bin, one wave
per bin), direct (per-pixel visibility words, no queues), tiled (per-chunk
tile lists, LDS visibility per tile) and queue (one persistent launch pulling
bin and tile-part items from a device queue; "queue" walks the columns of
every int32-path triangle, "queue-spans" covers compact ones by 8-row
spans)."""
import sys

import numpy as np
import pytest

from piko_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


MODES = ("binned", "direct", "tiled", "queue", "queue-spans")


def gpu_frame(scene, T, S, pair_capacity=0, mode="tiled"):
    import piko_amd
    rm = {"binned": piko_amd.RASTER_BINNED, "direct": piko_amd.RASTER_DIRECT, "tiled": piko_amd.RASTER_TILED,
          "queue": piko_amd.RASTER_QUEUE, "queue-spans": piko_amd.RASTER_QUEUE}[mode]
    cover = piko_amd.COVER_SPANS if mode == "queue-spans" else piko_amd.COVER_AUTO
    with piko_amd.RasterPipe().allocate(scene, T, S, pair_capacity, raster_mode=rm, queue_cover=cover) as pipe:
        pipe.prepare()
        pipe.run_single()
        return pipe.pikoScreen.getData(), pipe.get_depth(), pipe.get_bins(), pipe.stats()


def assert_frame_equal(scene, T, S, oracle_mod, check_bins=True, pair_capacity=0, modes=MODES):
    setup = oracle_mod.vertex_stage(scene, S)
    ostart, otris = oracle_mod.bins(setup, scene.W, scene.H, S, T)
    oc, od = oracle_mod.raster(setup, scene.W, scene.H, S, T, ostart, otris)
    for mode in modes:
        log("case W=%d H=%d n=%d T=%d S=%d kind=%s mode=%s" % (scene.W, scene.H, scene.n, T, S, scene.kind, mode))
        c, d, (bs, bt), st = gpu_frame(scene, T, S, pair_capacity, mode)
        assert st["emitted"] == int(setup.keep.sum())
        if check_bins:
            assert np.array_equal(bs, ostart), "bin starts differ"
            assert np.array_equal(bt, otris), "bin queues differ"
        bad = np.argwhere(c != oc)
        assert bad.size == 0, "%s: colour differs at %d pixels, first %s: gpu %08x oracle %08x" % (
            mode, len(bad), bad[0], c[tuple(bad[0])], oc[tuple(bad[0])])
        assert np.array_equal(d.view(np.int32), od.view(np.int32)), "%s: depth bits differ" % mode
    return c, d


@pytest.mark.parametrize("T", [8, 16, 32, 64])
@pytest.mark.parametrize("S", [4, 8])
def test_small_scene_all_bin_sizes(oracle_mod, T, S):
    sc = scenes.tris_uniform(3, 3000, 200, 150)
    assert_frame_equal(sc, T, S, oracle_mod)


@pytest.mark.parametrize("T", [8, 16, 32, 64])
def test_config1_full(oracle_mod, T):
    # 512x512, 50,000 triangles from seed 1, 8 subpixel bits
    sc = scenes.tris_uniform(1, 50000, 512, 512)
    assert_frame_equal(sc, T, 8, oracle_mod)


def test_large_triangles(oracle_mod):
    # triangles far larger than a bin, partly outside the screen and guard band
    rng = np.random.default_rng(4)
    W, H = 300, 170
    tris = []
    for _ in range(40):
        v = rng.uniform([-200, -200], [W + 200, H + 200], (3, 2))
        a = (v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[2, 0] - v[0, 0]) * (v[1, 1] - v[0, 1])
        if a < 0:
            v = v[[0, 2, 1]]
        tris.append(v)
    pos = np.zeros((len(tris), 9), np.float32)
    for i, v in enumerate(tris):
        pos[i, 0::3] = v[:, 0]
        pos[i, 1::3] = v[:, 1]
        pos[i, 2::3] = rng.uniform(0.1, 1.0, 3)
    col = rng.integers(0, 2**32, (len(tris), 3), dtype=np.uint64).astype(np.uint32)
    sc = scenes.TriScene(pos, col, W, H)
    for T in (8, 64):
        assert_frame_equal(sc, T, 8, oracle_mod)


def test_ties_and_full_screen(oracle_mod):
    W, H = 130, 70
    big = [(-10.0, -10.0), (300.0, -10.0), (-10.0, 250.0)]  # inside the 225-px guard band
    pos = np.zeros((4, 9), np.float32)
    for i, z in enumerate([0.5, 0.5, 0.25, 0.25]):
        pos[i, 0::3] = [p[0] for p in big]
        pos[i, 1::3] = [p[1] for p in big]
        pos[i, 2::3] = z
    col = np.array([[0xFF0000FF] * 3, [0xFF00FF00] * 3, [0xFFFF0000] * 3, [0xFF00FFFF] * 3], np.uint32)
    sc = scenes.TriScene(pos, col, W, H)
    c, d = assert_frame_equal(sc, 16, 8, oracle_mod)
    assert np.all(c == 0xFF00FFFF)  # last of the tied nearest triangles


@pytest.mark.parametrize("mode", MODES)
def test_empty_and_all_culled(oracle_mod, mode):
    sc = scenes.TriScene(np.zeros((0, 9), np.float32), np.zeros((0, 3), np.uint32), 64, 48)
    c, d, (bs, bt), st = gpu_frame(sc, 16, 8, mode=mode)
    assert np.all(c == 0xff663313) and np.all(d == 1.0) and st["pairs"] == 0 and len(bt) == 0
    cw = scenes.tris_uniform(2, 500, 64, 48)
    cw.pos = cw.pos[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]].copy()  # reverse winding: all back-facing
    c, d, (bs, bt), st = gpu_frame(cw, 16, 8, mode=mode)
    assert st["emitted"] == 0 and np.all(c == 0xff663313)


def sized_tris(seed, n, W, H, lo, hi):
    """n counter-clockwise triangles with bbox edges of lo..hi pixels, spread
    over the screen and a margin around it."""
    rng = np.random.default_rng(seed)
    c = rng.uniform([-0.1 * W, -0.1 * H], [1.1 * W, 1.1 * H], (n, 2))
    e = rng.uniform(lo, hi, (n, 1))
    ang = rng.uniform(0, 2 * np.pi, (n, 1)) + np.array([0.0, 2.1, 4.2]) + rng.uniform(-0.4, 0.4, (n, 3))
    pos = np.zeros((n, 9), np.float32)
    pos[:, 0::3] = c[:, :1] + 0.6 * e * np.cos(ang)
    pos[:, 1::3] = c[:, 1:] + 0.6 * e * np.sin(ang)
    pos[:, 2::3] = rng.uniform(0.05, 0.95, (n, 3))
    col = rng.integers(0, 2**32, (n, 3), dtype=np.uint64).astype(np.uint32)
    return scenes.TriScene(pos, col, W, H)


def test_direct_size_threshold(oracle_mod):
    # bboxes from a few pixels to ~190x190: both sides of the 16384-pixel
    # limit of the direct vertex stage, the rest drawn in 8x8 tiles
    sc = sized_tris(11, 1500, 400, 300, 4, 190)
    assert_frame_equal(sc, 16, 8, oracle_mod)


def test_direct_large_list_grows(oracle_mod):
    # more than 4096 large triangles: the large-triangle list overflows,
    # grows and the frame is re-run
    sc = sized_tris(12, 9000, 512, 384, 170, 200)
    assert_frame_equal(sc, 32, 8, oracle_mod, modes=("direct",))
    import piko_amd
    with piko_amd.RasterPipe().allocate(sc, 32, 8, raster_mode=piko_amd.RASTER_DIRECT) as pipe:
        pipe.run_frames(3)
        a = pipe.pikoScreen.getData().copy()
        pipe.run_single()
        assert np.array_equal(a, pipe.pikoScreen.getData())


def test_direct_large_after_frames_without(oracle_mod):
    # after a frame without large triangles the next frame draws any large
    # ones inside the vertex stage (no large-triangle launch); a camera that
    # brings large triangles in must give the oracle's image on that frame
    # and on the next (launched again)
    import piko_amd
    far = scenes.heightfield(seed=5, grid=60, W=320, H=200, eye=(0.0, 2.5, 3.0), target=(0.0, 0.0, 0.0))
    near = scenes.heightfield(seed=5, grid=60, W=320, H=200, eye=(0.0, 0.12, 0.25), target=(0.0, 0.0, -0.3))
    fs, ns = oracle_mod.vertex_stage(far, 8), oracle_mod.vertex_stage(near, 8)
    big = lambda st: int(((np.ptp(st.xy[st.keep.astype(bool)][:, 0::2], 1) >> 8) *
                          (np.ptp(st.xy[st.keep.astype(bool)][:, 1::2], 1) >> 8) > 1024).sum())
    assert big(fs) == 0 and big(ns) > 0
    oc, od, _ = oracle_mod.frame(near, 8, 16)
    with piko_amd.RasterPipe().allocate(far, 16, 8, raster_mode=piko_amd.RASTER_DIRECT) as pipe:
        pipe.prepare()
        pipe.run_single()
        for rep in range(2):
            log("near camera, frame %d" % rep)
            pipe.prepare(near)
            pipe.run_single()
            c, d = pipe.pikoScreen.getData(), pipe.get_depth()
            assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32)), rep


def test_direct_frames_across_camera_changes(oracle_mod):
    # runs of back-to-back direct frames with the camera changed between
    # runs, large triangles appearing and disappearing (drawn inline after a
    # run without any, then by the large-triangle launch): each run ends on
    # the oracle's image
    import piko_amd
    far = scenes.heightfield(seed=6, grid=80, W=400, H=240, eye=(0.0, 2.5, 3.0), target=(0.0, 0.0, 0.0))
    near = scenes.heightfield(seed=6, grid=80, W=400, H=240, eye=(0.0, 0.12, 0.25), target=(0.0, 0.0, -0.3))
    ofar, odfar, _ = oracle_mod.frame(far, 8, 16)
    onear, odnear, _ = oracle_mod.frame(near, 8, 16)
    with piko_amd.RasterPipe().allocate(far, 16, 8, raster_mode=piko_amd.RASTER_DIRECT) as pipe:
        pipe.prepare()
        for step, (sc, oc, od, n) in enumerate([(far, ofar, odfar, 4), (near, onear, odnear, 2), (near, onear, odnear, 5),
                                                (far, ofar, odfar, 3), (near, onear, odnear, 1)]):
            log("run %d: %d frames" % (step, n))
            pipe.prepare(sc)
            pipe.run_frames(n)
            c, d = pipe.pikoScreen.getData(), pipe.get_depth()
            assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32)), step


@pytest.mark.parametrize("cover", ["walk", "spans"])
@pytest.mark.parametrize("T", [32, 64])
def test_queue_frames_across_camera_changes(oracle_mod, T, cover):
    # queue-path runs of several frames with the camera changed between runs
    # and small parts (many split tiles): moving near the terrain overflows
    # tile ranges sized on the far view (the ranges grow and the frames
    # re-run), and each run ends on the oracle's image
    import piko_amd
    far = scenes.heightfield(seed=6, grid=80, W=400, H=240, eye=(0.0, 2.5, 3.0), target=(0.0, 0.0, 0.0))
    near = scenes.heightfield(seed=6, grid=80, W=400, H=240, eye=(0.0, 0.12, 0.25), target=(0.0, 0.0, -0.3))
    ofar, odfar, _ = oracle_mod.frame(far, 8, T)
    onear, odnear, _ = oracle_mod.frame(near, 8, T)
    qc = piko_amd.COVER_WALK if cover == "walk" else piko_amd.COVER_SPANS
    with piko_amd.RasterPipe().allocate(far, T, 8, raster_mode=piko_amd.RASTER_QUEUE, part_pairs=64,
                                        queue_cover=qc) as pipe:
        pipe.prepare()
        for step, (sc, oc, od, n) in enumerate([(far, ofar, odfar, 4), (near, onear, odnear, 2), (near, onear, odnear, 5),
                                                (far, ofar, odfar, 3), (near, onear, odnear, 1), (far, ofar, odfar, 6)]):
            log("run %d: %d frames" % (step, n))
            pipe.prepare(sc)
            pipe.run_frames(n)
            c, d = pipe.pikoScreen.getData(), pipe.get_depth()
            assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32)), step


@pytest.mark.parametrize("wpc", [1, 8, 64])
def test_queue_waves_per_cu(oracle_mod, wpc):
    # the persistent raster grid at one workgroup per CU, at two, and at a
    # request above what a CU holds (capped to the resident workgroups), with
    # small parts so that most workgroups take many items: the same image
    import piko_amd
    sc = scenes.heightfield(seed=4, grid=120, W=640, H=360)
    oc, od, _ = oracle_mod.frame(sc, 8, 32)
    log("queue waves_per_cu=%d" % wpc)
    with piko_amd.RasterPipe().allocate(sc, 32, 8, raster_mode=piko_amd.RASTER_QUEUE, waves_per_cu=wpc,
                                        part_pairs=128) as pipe:
        pipe.prepare()
        pipe.run_frames(2)
        c, d = pipe.pikoScreen.getData(), pipe.get_depth()
        assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32))


@pytest.mark.parametrize("nframes", [2, 3])
def test_queue_frames_in_flight_pretransformed(oracle_mod, nframes):
    # run_frames keeps two frames in flight (the pipe and its twin); with an
    # even count the last frame runs on the twin and its image is copied back
    import piko_amd
    sc = scenes.tris_uniform(21, 40000, 800, 600)
    oc, od, _ = oracle_mod.frame(sc, 8, 32)
    log("queue frames in flight: %d frames" % nframes)
    with piko_amd.RasterPipe().allocate(sc, 32, 8, raster_mode=piko_amd.RASTER_QUEUE) as pipe:
        pipe.prepare()
        pipe.run_frames(nframes)
        c, d = pipe.pikoScreen.getData(), pipe.get_depth()
        assert np.array_equal(c, oc) and np.array_equal(d.view(np.int32), od.view(np.int32))
        st = pipe.stats()
        pipe.run_single()
        assert pipe.stats()["emitted"] == st["emitted"]
        assert np.array_equal(pipe.pikoScreen.getData(), oc)


def test_tiled_large_lists(oracle_mod):
    # every triangle spans more than 2x2 tiles of 16 pixels: all of them go to
    # the chunks' large lists and are clipped to each tile they touch
    sc = sized_tris(13, 9000, 512, 384, 45, 90)
    assert_frame_equal(sc, 16, 8, oracle_mod, check_bins=False, modes=("tiled",))


def test_tiled_tile_size_grows_with_screen(oracle_mod):
    # 2000x1200 at bin size 8 or 16 would be more than 8192 tiles: the tiled
    # path rasterises it in 32-pixel tiles, with the same image
    sc = scenes.tris_uniform(14, 20000, 2000, 1200)
    assert_frame_equal(sc, 8, 8, oracle_mod, check_bins=False, modes=("tiled",))


def test_queue_capacity_grows(oracle_mod):
    sc = scenes.tris_uniform(8, 20000, 256, 256)
    assert_frame_equal(sc, 8, 8, oracle_mod, pair_capacity=1000)


def test_world_heightfield_small(oracle_mod):
    sc = scenes.heightfield(seed=2, grid=200, W=480, H=270)
    for T in (8, 16, 32, 64):
        assert_frame_equal(sc, T, 8, oracle_mod)


@pytest.mark.parametrize("T", [8, 16, 32, 64])
def test_config2_heightfield_full(oracle_mod, T):
    # 1920x1080, 2,000,000-triangle height field (depth complexity 3.94),
    # Lambert vertex lighting, every path at every bin size
    sc = scenes.heightfield()
    assert_frame_equal(sc, T, 8, oracle_mod, check_bins=(T in (8, 32)))


def test_deterministic_repeat():
    import piko_amd
    sc = scenes.tris_uniform(1, 50000, 512, 512)
    with piko_amd.RasterPipe().allocate(sc, 16, 8) as pipe:
        pipe.run_single()
        a = pipe.pikoScreen.getData().copy()
        pipe.run_frames(5)
        b = pipe.pikoScreen.getData()
    assert np.array_equal(a, b)
