"""The C-ABI libraries load without a GPU and export every function the
headers under include/ declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "piko-public_amd", "piko_amd")


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(piko_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("piko_hip.h", "libpiko_hip.so"), ("piko_scene.h", "libpiko_scene.so")])
def test_library_exports_header(header, lib):
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) > 3
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_vertex_module_holds_pipeline_names():
    """The code object the runtime shim loads (piko_rt_module_load) carries
    the module global and entry names the generated host code fetches."""
    path = os.path.join(PKG, "vertex_module.hsaco")
    assert os.path.exists(path), "build the code object first (__graft_entry__.build())"
    data = open(path, "rb").read()
    assert b"amdgcn-amd-amdhsa" in data and b"gfx950" in data
    for name in (b"constState", b"kernel0", b"kernel1"):
        assert name in data, name


def test_status_strings():
    L = ctypes.CDLL(os.path.join(PKG, "libpiko_hip.so"))
    L.piko_status_string.restype = ctypes.c_char_p
    assert L.piko_status_string(0) == b"ok"
    assert L.piko_status_string(-4) == b"queue capacity exceeded"


def test_invalid_arguments_rejected_without_device():
    """Argument validation runs before any device call."""
    import sys
    sys.path.insert(0, os.path.join(REPO, "piko-public_amd"))
    import piko_amd
    L = piko_amd.lib()
    h = ctypes.c_void_p()
    cs = piko_amd.ConstantState()
    cs.screenSizeX, cs.screenSizeY = 64, 64
    bad_bin = piko_amd.Options(12, 8, 0, 0)
    assert L.piko_raster_allocate(ctypes.byref(h), ctypes.byref(cs), None, None, 0, ctypes.byref(bad_bin)) == -1
    bad_s = piko_amd.Options(16, 9, 0, 0)
    assert L.piko_raster_allocate(ctypes.byref(h), ctypes.byref(cs), None, None, 0, ctypes.byref(bad_s)) == -1
    bad_cover = piko_amd.Options(16, 8, 0, 0, 0, piko_amd.RASTER_QUEUE, 0, 3)
    assert L.piko_raster_allocate(ctypes.byref(h), ctypes.byref(cs), None, None, 0, ctypes.byref(bad_cover)) == -1
    bo = piko_amd.BatchOptions(0, 8, 4, 1, 4096, 0, piko_amd.RASTER_QUEUE, 0, 0, -1)
    assert L.piko_batch_create(ctypes.byref(h), None, None, 0, 0, ctypes.byref(bo)) == -1
    assert L.piko_raster_run_single(None) == -1


def test_scene_generators_deterministic():
    import sys
    sys.path.insert(0, os.path.join(REPO, "piko-public_amd"))
    import numpy as np
    from piko_amd import scenes
    a = scenes.tris_uniform(1, 1000, 512, 512)
    b = scenes.tris_uniform(1, 1000, 512, 512)
    c = scenes.tris_uniform(2, 1000, 512, 512)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.col, b.col)
    assert not np.array_equal(a.pos, c.pos)
    # counter-based: a prefix of a scene is the smaller scene
    d = scenes.tris_uniform(1, 100, 512, 512)
    assert np.array_equal(a.pos[:100], d.pos)
    hf = scenes.heightfield(grid=20, W=64, H=36)
    assert hf.n == 800
    n = hf.nrm.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    spec = scenes.batch_spec()
    assert len(spec) == 256 and spec[0][0] == 1000 and spec[-1][0] == 1255
    ns = np.array([s[1] for s in spec])
    px = np.array([s[2] * s[3] for s in spec])
    assert ns.min() >= 1000 and ns.max() <= 500000
    assert px.min() >= 0.95 * 256 ** 2 and px.max() <= 1.05 * 2048 ** 2


def test_reyes_global_sampling_screen_bound_without_device():
    """Global sampling scans a wave's (micropolygon, pixel) pairs in int32,
    so it is refused where 64*W*H reaches 2^31; AUTO falls to tiled sampling
    there, and a screen neither path can hold is refused."""
    import sys
    sys.path.insert(0, os.path.join(REPO, "piko-public_amd"))
    import piko_amd
    L = piko_amd.lib()
    L.piko_reyes_allocate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(piko_amd.ConstantState),
                                      ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(piko_amd.ReyesOptions)]
    h = ctypes.c_void_p()

    def alloc(W, H, sampling):
        cs = piko_amd.ConstantState()
        cs.screenSizeX, cs.screenSizeY = W, H
        opt = piko_amd.ReyesOptions(16.0, 20, 8, 4, 0.0, 0, 0, sampling, 0)
        return L.piko_reyes_allocate(ctypes.byref(h), ctypes.byref(cs), None, 0, ctypes.byref(opt))

    # 8192 x 4096: 64*W*H = 2^31 exactly -> global refused; 32-px tiles: 256 x 128 > 8192 tiles
    assert alloc(8192, 4096, piko_amd.SAMPLE_GLOBAL) == -1
    assert alloc(8192, 4096, piko_amd.SAMPLE_AUTO) == -1
    # without a device every admissible request stops at the device check (-3)
    ok = alloc(4096, 4095, piko_amd.SAMPLE_GLOBAL)
    assert ok in (-3, 0)
    if ok == 0:
        L.piko_reyes_destroy(h)
