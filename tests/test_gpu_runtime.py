"""GPU: the runtime shim (piko_rt_*) and the device-input allocators, through
the C-ABI.

* piko_rt_module_load / get_function / get_global / launch on a gfx950 code
  object built here (piko_amd/vertex_module.hsaco): the uniforms go into the
  module global constState (the order of the reference's generated host code,
  PTXBackend.cpp:242-246), kernel0 runs the world vertex stage and its
  screen vertices, depths, lit colours and keep flags equal the oracle's
  vertex stage bit for bit; kernel1 clears the screen from constState; events
  time a launch on a created stream; missing names and files are refused.
* piko_raster_allocate_device and piko_reyes_allocate_device render from
  buffers already in device memory and equal the oracle's frames.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import piko_amd
from piko_amd import scenes

pytestmark = pytest.mark.gpu

MODULE = os.path.join(os.path.dirname(piko_amd.lib_path()), "vertex_module.hsaco")
P, I, SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rt():
    L = piko_amd.lib()
    L.piko_rt_init.argtypes = [I]
    L.piko_rt_module_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    L.piko_rt_module_unload.argtypes = [P]
    L.piko_rt_get_function.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(P)]
    L.piko_rt_get_global.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(SZ)]
    L.piko_rt_malloc.argtypes = [ctypes.POINTER(P), SZ]
    L.piko_rt_free.argtypes = [P]
    L.piko_rt_memcpy_htod.argtypes = [P, P, SZ]
    L.piko_rt_memcpy_dtoh.argtypes = [P, P, SZ]
    L.piko_rt_memset_d32.argtypes = [P, ctypes.c_uint32, SZ]
    L.piko_rt_launch.argtypes = [P, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, P, ctypes.POINTER(P)]
    L.piko_rt_sync.argtypes = [P]
    L.piko_rt_stream_create.argtypes = [ctypes.POINTER(P)]
    L.piko_rt_stream_destroy.argtypes = [P]
    L.piko_rt_event_create.argtypes = [ctypes.POINTER(P)]
    L.piko_rt_event_destroy.argtypes = [P]
    L.piko_rt_event_record.argtypes = [P, P]
    L.piko_rt_event_elapsed.argtypes = [P, P, ctypes.POINTER(ctypes.c_float)]
    L.piko_rt_mem_info.argtypes = [ctypes.POINTER(SZ), ctypes.POINTER(SZ)]
    L.piko_rt_error_string.restype = ctypes.c_char_p
    return L


def ok(rc, what):
    assert rc == 0, "%s: %s" % (what, piko_amd.STATUS.get(rc, rc))


class Dev:
    """A device buffer of n elements of dtype, owned by the test."""

    def __init__(self, L, n, dtype):
        self.L, self.n, self.dtype = L, n, np.dtype(dtype)
        self.ptr = P()
        ok(L.piko_rt_malloc(ctypes.byref(self.ptr), max(n * self.dtype.itemsize, 4)), "rt_malloc")

    def put(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        ok(self.L.piko_rt_memcpy_htod(self.ptr, a.ctypes.data_as(P), a.nbytes), "rt_memcpy_htod")

    def get(self):
        out = np.empty(self.n, self.dtype)
        ok(self.L.piko_rt_memcpy_dtoh(out.ctypes.data_as(P), self.ptr, out.nbytes), "rt_memcpy_dtoh")
        return out

    def free(self):
        self.L.piko_rt_free(self.ptr)


def args_array(*values):
    """The kernelParams array of cuLaunchKernel: one pointer per argument."""
    arr = (P * len(values))()
    for k, v in enumerate(values):
        arr[k] = ctypes.cast(ctypes.byref(v), P)
    return arr


@pytest.fixture(scope="module")
def module():
    if not os.path.exists(MODULE):
        pytest.fail("vertex_module.hsaco is not built (run __graft_entry__.build())")
    L = rt()
    ok(L.piko_rt_init(0), "rt_init")
    mod = P()
    ok(L.piko_rt_module_load(MODULE.encode(), ctypes.byref(mod)), "rt_module_load")
    yield L, mod
    ok(L.piko_rt_module_unload(mod), "rt_module_unload")


def set_const_state(L, mod, cs):
    d, nbytes = P(), SZ()
    ok(L.piko_rt_get_global(mod, b"constState", ctypes.byref(d), ctypes.byref(nbytes)), "rt_get_global")
    assert nbytes.value == ctypes.sizeof(piko_amd.ConstantState)
    ok(L.piko_rt_memcpy_htod(d, ctypes.cast(ctypes.byref(cs), P), ctypes.sizeof(cs)), "constState upload")
    back = piko_amd.ConstantState()
    ok(L.piko_rt_memcpy_dtoh(ctypes.cast(ctypes.byref(back), P), d, ctypes.sizeof(back)), "constState read")
    assert bytes(back) == bytes(cs)


@pytest.mark.parametrize("case", ["heightfield", "ragged"])
def test_module_vertex_stage_equals_oracle(module, oracle_mod, case):
    L, mod = module
    if case == "heightfield":
        scene = scenes.heightfield(seed=2, grid=200, W=480, H=270)
    else:  # a count that is not a multiple of the workgroup, on another screen
        scene = scenes.heightfield(seed=5, grid=37, W=333, H=191)
    n, S = scene.n, 8
    log("module kernel0 %s: %d triangles %dx%d" % (case, n, scene.W, scene.H))
    set_const_state(L, mod, piko_amd.ConstantState.for_scene(scene))
    fn = P()
    ok(L.piko_rt_get_function(mod, b"kernel0", ctypes.byref(fn)), "rt_get_function kernel0")
    pos, nrm = Dev(L, 9 * n, np.float32), Dev(L, 9 * n, np.float32)
    xy, z = Dev(L, 6 * n, np.int32), Dev(L, 3 * n, np.float32)
    col, keep = Dev(L, 3 * n, np.uint32), Dev(L, n, np.uint8)
    try:
        pos.put(scene.pos.reshape(-1))
        nrm.put(scene.nrm.reshape(-1))
        keep.put(np.full(n, 7, np.uint8))
        a = args_array(pos.ptr, nrm.ptr, I(n), I(S), xy.ptr, z.ptr, col.ptr, keep.ptr)
        ok(L.piko_rt_launch(fn, (n + 255) // 256, 256, 0, None, a), "rt_launch kernel0")
        ok(L.piko_rt_sync(None), "rt_sync")
        setup = oracle_mod.vertex_stage(scene, S)
        assert np.array_equal(xy.get(), setup.xy.reshape(-1)), "screen vertices differ"
        assert np.array_equal(z.get().view(np.int32), setup.z.reshape(-1).view(np.int32)), "depths differ"
        assert np.array_equal(col.get(), setup.col.reshape(-1)), "vertex colours differ"
        assert np.array_equal(keep.get(), setup.keep), "keep flags differ"
        assert 0 < int(setup.keep.sum()) < n
    finally:
        for b in (pos, nrm, xy, z, col, keep):
            b.free()


def test_module_clear_on_stream_with_events(module):
    L, mod = module
    W, H = 300, 170
    scene = scenes.tris_uniform(1, 10, W, H)
    set_const_state(L, mod, piko_amd.ConstantState.for_scene(scene))
    fn = P()
    ok(L.piko_rt_get_function(mod, b"kernel1", ctypes.byref(fn)), "rt_get_function kernel1")
    color, depth = Dev(L, W * H + 64, np.uint32), Dev(L, W * H + 64, np.float32)
    stream, e0, e1 = P(), P(), P()
    ok(L.piko_rt_stream_create(ctypes.byref(stream)), "rt_stream_create")
    ok(L.piko_rt_event_create(ctypes.byref(e0)), "rt_event_create")
    ok(L.piko_rt_event_create(ctypes.byref(e1)), "rt_event_create")
    try:
        # the 64 words past W*H are a guard: kernel1 must stop at W*H
        ok(L.piko_rt_memset_d32(color.ptr, 0x12345678, W * H + 64), "rt_memset_d32")
        ok(L.piko_rt_memset_d32(depth.ptr, 0x40000000, W * H + 64), "rt_memset_d32")
        ok(L.piko_rt_sync(None), "rt_sync")
        ok(L.piko_rt_event_record(e0, stream), "rt_event_record")
        a = args_array(color.ptr, depth.ptr)
        ok(L.piko_rt_launch(fn, (W * H + 255) // 256, 256, 0, stream, a), "rt_launch kernel1")
        ok(L.piko_rt_event_record(e1, stream), "rt_event_record")
        ok(L.piko_rt_sync(stream), "rt_sync stream")
        ms = ctypes.c_float()
        ok(L.piko_rt_event_elapsed(e0, e1, ctypes.byref(ms)), "rt_event_elapsed")
        assert 0.0 <= ms.value < 1000.0
        c, d = color.get(), depth.get()
        assert (c[:W * H] == 0xFF663313).all() and (c[W * H:] == 0x12345678).all()
        assert (d[:W * H] == 1.0).all() and (d[W * H:] == 2.0).all()
        free_b, total_b = SZ(), SZ()
        ok(L.piko_rt_mem_info(ctypes.byref(free_b), ctypes.byref(total_b)), "rt_mem_info")
        assert 0 < free_b.value <= total_b.value
    finally:
        L.piko_rt_event_destroy(e0)
        L.piko_rt_event_destroy(e1)
        L.piko_rt_stream_destroy(stream)
        color.free()
        depth.free()


def test_module_missing_names_and_files_refused(module, tmp_path):
    L, mod = module
    fn, d, nbytes = P(), P(), SZ()
    assert L.piko_rt_get_function(mod, b"kernel7", ctypes.byref(fn)) != 0
    assert L.piko_rt_get_global(mod, b"mutableState", ctypes.byref(d), ctypes.byref(nbytes)) != 0
    m2 = P()
    assert L.piko_rt_module_load(str(tmp_path / "absent.hsaco").encode(), ctypes.byref(m2)) != 0
    assert L.piko_rt_get_function(None, b"kernel0", ctypes.byref(fn)) != 0
    assert L.piko_rt_launch(None, 1, 64, 0, None, None) != 0


@pytest.mark.parametrize("mode", ["direct", "tiled", "queue"])
def test_raster_allocate_device_equals_oracle(oracle_mod, mode):
    scene = scenes.heightfield(seed=2, grid=200, W=480, H=270)
    T = 32
    log("raster allocate_device %s: %d triangles %dx%d T=%d" % (mode, scene.n, scene.W, scene.H, T))
    cs = piko_amd.ConstantState.for_scene(scene)
    dpos, dnrm = piko_amd.DeviceArray(scene.pos), piko_amd.DeviceArray(scene.nrm)
    raster_mode = {"direct": piko_amd.RASTER_DIRECT, "tiled": piko_amd.RASTER_TILED,
                   "queue": piko_amd.RASTER_QUEUE}[mode]
    try:
        with piko_amd.RasterPipe().allocate_device(dpos.ptr, dnrm.ptr, scene.n, cs, kind="world", bin_size=T,
                                                   subpixel_bits=8, raster_mode=raster_mode) as pipe:
            for _ in range(2):  # the second frame reuses every buffer of the first
                pipe.run_single()
                color, depth = pipe.pikoScreen.getData(), pipe.get_depth()
                oc, od = oracle_mod.frame(scene, 8, T)[:2]
                assert np.array_equal(color, oc), "colour differs"
                assert np.array_equal(depth.view(np.int32), od.view(np.int32)), "depth differs"
    finally:
        dpos.free()
        dnrm.free()


def test_reyes_allocate_device_equals_oracle(oracle_mod):
    scene = scenes.sphere_patches(seed=3, nlat=8, nlon=8, W=320, H=180, spp=4, disp=0.05)
    log("reyes allocate_device: %d patches %dx%d" % (scene.n, scene.W, scene.H))
    orgba, oz, nup = oracle_mod.reyes_frame(scene)
    dcp = piko_amd.DeviceArray(np.ascontiguousarray(scene.cp, np.float32))
    try:
        for sampling, tile in ((piko_amd.SAMPLE_GLOBAL, 0), (piko_amd.SAMPLE_TILED, 16)):
            with piko_amd.ReyesPipe().allocate_device(scene, dcp.ptr, scene.n, sampling=sampling,
                                                      tile=tile) as pipe:
                pipe.prepare()
                pipe.run_single()
                rgba, depth, st = pipe.color(), pipe.depth(), pipe.stats()
            assert st["micropolygons"] == nup
            assert np.array_equal(rgba.view(np.int32), orgba.view(np.int32)), "colour differs"
            assert np.array_equal(depth.view(np.int32), oz.min(axis=2)), "depth differs"
    finally:
        dcp.free()
