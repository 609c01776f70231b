"""The allocation-free frame (piko_raster_frame_noalloc, no device
allocation, per-frame buffers in the library's fixed scratch) against the
pipe's frame on the same device-resident input, through the C-ABI: colour
words and depth bits equal, two calls in a row with the second on a smaller
input, then a world-space scene, then sizes beyond the scratch refused."""
import ctypes
import sys

import numpy as np
import pytest

import piko_amd
from piko_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def read_back(dev, n, dtype):
    L = piko_amd.lib()
    L.piko_rt_memcpy_dtoh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    out = np.empty(n, dtype)
    assert L.piko_rt_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev.ptr), out.nbytes) == 0
    return out


def both_frames(scene, S=8):
    attr = scene.col if scene.kind == "pretransformed" else scene.nrm
    dpos, dattr = piko_amd.DeviceArray(scene.pos), piko_amd.DeviceArray(attr)
    npix = scene.W * scene.H
    dcol = piko_amd.DeviceArray(np.full(npix, 0xDEADBEEF, np.uint32))
    ddep = piko_amd.DeviceArray(np.full(npix, -7.0, np.float32))
    cs = piko_amd.ConstantState.for_scene(scene)
    try:
        rc = piko_amd.frame_noalloc(cs, dpos.ptr, dattr.ptr, scene.n, scene.kind, S, dcol.ptr, ddep.ptr)
        assert rc == 0, piko_amd.STATUS.get(rc, rc)
        c = read_back(dcol, npix, np.uint32).reshape(scene.H, scene.W)
        d = read_back(ddep, npix, np.float32).reshape(scene.H, scene.W)
        with piko_amd.RasterPipe().allocate_device(dpos.ptr, dattr.ptr, scene.n, cs, kind=scene.kind, bin_size=16,
                                                   subpixel_bits=S, raster_mode=piko_amd.RASTER_DIRECT) as pipe:
            pipe.run_single()
            pc, pd = pipe.pikoScreen.getData(), pipe.get_depth()
    finally:
        for a in (dpos, dattr, dcol, ddep):
            a.free()
    return c, d, pc, pd


@pytest.mark.parametrize("case", ["config1", "smaller", "world", "large"])
def test_noalloc_frame_equals_pipe_frame(case):
    scene = {
        "config1": lambda: scenes.tris_uniform(1, 50000, 512, 512),
        "smaller": lambda: scenes.tris_uniform(2, 3000, 200, 150),
        "world": lambda: scenes.heightfield(seed=2, grid=200, W=480, H=270),
        # large triangles (bbox over 16,384 pixels) go through the large list
        "large": lambda: scenes.tris_uniform(3, 400, 640, 480, area_median=30000.0, sigma=0.3),
    }[case]()
    log("noalloc case %s: W=%d H=%d n=%d" % (case, scene.W, scene.H, scene.n))
    c, d, pc, pd = both_frames(scene)
    assert np.array_equal(c, pc), "colour differs at %d pixels" % int((c != pc).sum())
    assert np.array_equal(d.view(np.int32), pd.view(np.int32))


def test_noalloc_refuses_sizes_beyond_scratch():
    cs = piko_amd.ConstantState()
    cs.screenSizeX, cs.screenSizeY = 64, 64
    # more triangles than the scratch holds: refused before any device access
    assert piko_amd.frame_noalloc(cs, 16, 16, (1 << 20) + 1, "world", 8, 16, 16) == -7
    cs.screenSizeX, cs.screenSizeY = 2048, 2048  # 4,194,304 pixels > 2,097,152
    assert piko_amd.frame_noalloc(cs, 16, 16, 10, "world", 8, 16, 16) == -7
