"""Python host mirror of Piko's RasterPipe and ReyesPipe over the MI355X C-ABI
(libpiko_hip.so).

The reference's user surface is the generated C++ pipe
(pikoc/samples/rasterPipelineFixPt/rasterPipe.h:16-45):

    RasterPipe pipe;
    pipe.allocate(constState, mutableState, triangles, count);
    pipe.prepare(); pipe.run_single();
    unsigned* fb = pipe.pikoScreen.getData();
    pipe.destroy();

This module keeps those names and argument meanings.  Every call goes to the
HIP library; there is no CPU path: if libpiko_hip.so is missing, or no GPU is
present, the call raises.
"""
import ctypes
import os

import numpy as np

from . import scenes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

PIKO_INPUT_PRETRANSFORMED = 0
PIKO_INPUT_WORLD = 1
RASTER_AUTO, RASTER_BINNED, RASTER_DIRECT, RASTER_TILED, RASTER_QUEUE = 0, 1, 2, 3, 4  # PIKO_RASTER_* (include/piko_hip.h)
COVER_AUTO, COVER_WALK, COVER_SPANS = 0, 1, 2  # PIKO_COVER_* (include/piko_hip.h)
SAMPLE_AUTO, SAMPLE_GLOBAL, SAMPLE_TILED = 0, 1, 2  # PIKO_SAMPLE_* (include/piko_hip.h)

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "no HIP device",
          -4: "queue capacity exceeded", -5: "device out of memory", -6: "admission refused",
          -7: "larger than the fixed scratch of the allocation-free frame"}


class PikoError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s: %s (%d)" % (what, STATUS.get(status, "unknown"), status))


class ConstantState(ctypes.Structure):
    """piko_constant_state (include/piko_hip.h); ConstantState of
    api/include/piko/builtinTypes.h:20-33 plus the lighting constants."""
    _fields_ = [("screenSizeX", ctypes.c_int), ("screenSizeY", ctypes.c_int),
                ("halfW", ctypes.c_float), ("halfH", ctypes.c_float),
                ("viewProjMatrix", ctypes.c_float * 16),
                ("light", ctypes.c_float * 3), ("matcol", ctypes.c_float * 3),
                ("ambient", ctypes.c_float * 3)]

    @classmethod
    def for_scene(cls, scene):
        cs = cls()
        cs.screenSizeX, cs.screenSizeY = scene.W, scene.H
        # initScreen, rasterPipelineFixPt/main.cpp:104-105
        cs.halfW = np.float32(0.5) * np.float32(scene.W)
        cs.halfH = np.float32(0.5) * np.float32(scene.H)
        vp = getattr(scene, "view_proj", None)
        if vp is not None:
            for k in range(16):
                cs.viewProjMatrix[k] = float(vp[k])
        for name, default in (("light", scenes.DEFAULT_LIGHT), ("matcol", scenes.DEFAULT_MATCOL),
                              ("ambient", scenes.DEFAULT_AMBIENT)):
            v = getattr(scene, name, None)
            v = default if v is None else v
            arr = getattr(cs, name)
            for k in range(3):
                arr[k] = float(np.float32(v[k]))
        return cs


class Options(ctypes.Structure):
    _fields_ = [("bin_size", ctypes.c_int), ("subpixel_bits", ctypes.c_int),
                ("input_kind", ctypes.c_int), ("pair_capacity", ctypes.c_longlong),
                ("waves_per_cu", ctypes.c_int), ("raster_mode", ctypes.c_int), ("part_pairs", ctypes.c_int),
                ("queue_cover", ctypes.c_int)]


class RasterJob(ctypes.Structure):
    """piko_raster_job: one screen of a batch."""
    _fields_ = [("count", ctypes.c_int), ("src_base", ctypes.c_longlong), ("bin_size", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("queue_depth", ctypes.c_float), ("cs", ConstantState),
                ("queue_cover", ctypes.c_int)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("input_kind", ctypes.c_int), ("subpixel_bits", ctypes.c_int), ("max_jobs", ctypes.c_int),
                ("max_tris", ctypes.c_longlong), ("max_pixels", ctypes.c_longlong),
                ("pair_capacity", ctypes.c_longlong), ("raster_mode", ctypes.c_int), ("part_pairs", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("queue_cover", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_longlong), ("emitted", ctypes.c_longlong),
                ("pairs", ctypes.c_longlong), ("pair_capacity", ctypes.c_longlong),
                ("bins_x", ctypes.c_int), ("bins_y", ctypes.c_int),
                ("device_bytes", ctypes.c_longlong)]


class ReyesOptions(ctypes.Structure):
    """piko_reyes_options (include/piko_hip.h)."""
    _fields_ = [("threshold", ctypes.c_float), ("max_splits", ctypes.c_int), ("dice", ctypes.c_int),
                ("spp", ctypes.c_int), ("displacement", ctypes.c_float), ("roots_per_wave", ctypes.c_int),
                ("keep_samples", ctypes.c_int), ("sampling", ctypes.c_int), ("tile", ctypes.c_int)]


class ReyesStats(ctypes.Structure):
    _fields_ = [("patches", ctypes.c_longlong), ("leaves", ctypes.c_longlong),
                ("micropolygons", ctypes.c_longlong), ("levels", ctypes.c_int),
                ("device_bytes", ctypes.c_longlong), ("capacity_reruns", ctypes.c_int),
                ("tile_pairs", ctypes.c_longlong)]


def lib_path():
    return os.path.join(_HERE, "libpiko_hip.so")


def lib():
    """Loads libpiko_hip.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libpiko_hip.so is not built (run __graft_entry__.build()); "
                               "the raster path has no CPU fallback")
        L = ctypes.CDLL(path)
        P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.piko_status_string.restype = ctypes.c_char_p
        L.piko_raster_allocate.argtypes = [pp, ctypes.POINTER(ConstantState), P, P, I, ctypes.POINTER(Options)]
        L.piko_raster_allocate_device.argtypes = [pp, ctypes.POINTER(ConstantState), P, P, I,
                                                  ctypes.POINTER(Options)]
        L.piko_raster_prepare.argtypes = [P, ctypes.POINTER(ConstantState)]
        L.piko_raster_run_single.argtypes = [P]
        L.piko_raster_run_frames.argtypes = [P, I]
        L.piko_raster_profile_frames.argtypes = [P, I, P, P, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_float)]
        L.piko_raster_get_data.argtypes = [P, P]
        L.piko_raster_get_depth.argtypes = [P, P]
        L.piko_raster_get_bins.argtypes = [P, P, P, LL]
        L.piko_raster_get_bins.restype = LL
        L.piko_raster_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        L.piko_raster_destroy.argtypes = [P]
        L.piko_batch_create.argtypes = [pp, P, P, LL, I, ctypes.POINTER(BatchOptions)]
        L.piko_batch_submit.argtypes = [P, ctypes.POINTER(RasterJob), I]
        L.piko_batch_query.argtypes = [P]
        L.piko_batch_wait.argtypes = [P]
        L.piko_batch_get_color.argtypes = [P, I, P]
        L.piko_batch_get_depth.argtypes = [P, I, P]
        L.piko_reyes_allocate.argtypes = [pp, ctypes.POINTER(ConstantState), P, I, ctypes.POINTER(ReyesOptions)]
        L.piko_reyes_allocate_device.argtypes = [pp, ctypes.POINTER(ConstantState), P, I,
                                                 ctypes.POINTER(ReyesOptions)]
        L.piko_reyes_prepare.argtypes = [P, ctypes.POINTER(ConstantState)]
        L.piko_reyes_run_single.argtypes = [P]
        L.piko_reyes_run_frames.argtypes = [P, I]
        L.piko_reyes_profile_frames.argtypes = [P, I, P, P, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_float)]
        L.piko_reyes_get_color.argtypes = [P, P]
        L.piko_reyes_get_depth.argtypes = [P, P]
        L.piko_reyes_get_samples.argtypes = [P, P]
        L.piko_reyes_get_leaves.argtypes = [P, P, P, P, LL]
        L.piko_reyes_get_leaves.restype = LL
        L.piko_reyes_get_stats.argtypes = [P, ctypes.POINTER(ReyesStats)]
        L.piko_reyes_destroy.argtypes = [P]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise PikoError(rc, what)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class PikoScreen:
    """PikoScreen (api/include/piko/stage.h:53-136): the colour target."""

    def __init__(self, pipe):
        self._pipe = pipe

    def getData(self):
        """Colour buffer as uint32 ABGR8 [H, W], row 0 at the bottom (stage.h:108-117)."""
        p = self._pipe
        out = np.empty(p.W * p.H, np.uint32)
        _check(lib().piko_raster_get_data(p._h, _p(out)), "getData")
        return out.reshape(p.H, p.W)


class RasterPipe:
    """VertexShade -> Raster(bin T) -> depth -> Gouraud -> PikoScreen on the GPU."""

    def __init__(self):
        self._h = None
        self.pikoScreen = PikoScreen(self)

    def allocate(self, scene, bin_size=16, subpixel_bits=8, pair_capacity=0, waves_per_cu=0, raster_mode=0,
                 part_pairs=0, queue_cover=0):
        """RasterPipe::allocate (PTXBackend.cpp:649-789): uploads the input and
        sizes every stage.  `scene` is a scenes.TriScene or scenes.WorldScene.
        raster_mode: RASTER_AUTO, RASTER_BINNED, RASTER_DIRECT, RASTER_TILED or
        RASTER_QUEUE (same image); part_pairs: queue entries per part item and
        queue_cover: COVER_AUTO, COVER_WALK or COVER_SPANS (RASTER_QUEUE)."""
        if self._h is not None:
            raise RuntimeError("pipe already allocated")
        self.W, self.H, self.T, self.S = scene.W, scene.H, bin_size, subpixel_bits
        self.kind = PIKO_INPUT_PRETRANSFORMED if scene.kind == "pretransformed" else PIKO_INPUT_WORLD
        self.cs = ConstantState.for_scene(scene)
        opt = Options(bin_size, subpixel_bits, self.kind, pair_capacity, waves_per_cu, raster_mode, part_pairs,
                      queue_cover)
        pos = np.ascontiguousarray(scene.pos, np.float32)
        attr = np.ascontiguousarray(scene.col if self.kind == 0 else scene.nrm)
        h = ctypes.c_void_p()
        _check(lib().piko_raster_allocate(ctypes.byref(h), ctypes.byref(self.cs), _p(pos), _p(attr),
                                          int(scene.n), ctypes.byref(opt)), "allocate")
        self._h = h
        return self

    def allocate_device(self, d_pos, d_attr, count, cs, kind="world", bin_size=16, subpixel_bits=8,
                        pair_capacity=0, waves_per_cu=0, raster_mode=0, part_pairs=0, queue_cover=0):
        """RasterPipe::allocate over input that already lives on the device
        (piko_raster_allocate_device): d_pos / d_attr are device addresses of
        `count` triangles; cs is the ConstantState (screen size and camera)."""
        if self._h is not None:
            raise RuntimeError("pipe already allocated")
        self.W, self.H, self.T, self.S = cs.screenSizeX, cs.screenSizeY, bin_size, subpixel_bits
        self.kind = PIKO_INPUT_PRETRANSFORMED if kind == "pretransformed" else PIKO_INPUT_WORLD
        self.cs = cs
        opt = Options(bin_size, subpixel_bits, self.kind, pair_capacity, waves_per_cu, raster_mode, part_pairs,
                      queue_cover)
        h = ctypes.c_void_p()
        _check(lib().piko_raster_allocate_device(ctypes.byref(h), ctypes.byref(self.cs), ctypes.c_void_p(d_pos),
                                                 ctypes.c_void_p(d_attr), int(count), ctypes.byref(opt)),
               "allocate_device")
        self._h = h
        return self

    def prepare(self, scene=None, cs=None):
        """RasterPipe::prepare (PTXBackend.cpp:791-809): per-frame constant state
        (from a scene, or a ConstantState)."""
        if scene is not None:
            self.cs = ConstantState.for_scene(scene)
        if cs is not None:
            self.cs = cs
        _check(lib().piko_raster_prepare(self._h, ctypes.byref(self.cs)), "prepare")

    def run_single(self):
        _check(lib().piko_raster_run_single(self._h), "run_single")

    def run_frames(self, n):
        _check(lib().piko_raster_run_frames(self._h, int(n)), "run_frames")

    def profile_frames(self, n, cap=64):
        ms = (ctypes.c_float * cap)()
        names = (ctypes.c_char_p * cap)()
        nk = ctypes.c_int()
        fm = ctypes.c_float()
        _check(lib().piko_raster_profile_frames(self._h, int(n), ctypes.cast(ms, ctypes.c_void_p),
                                                ctypes.cast(names, ctypes.c_void_p), cap, ctypes.byref(nk),
                                                ctypes.byref(fm)), "profile_frames")
        return [(names[k].decode(), float(ms[k])) for k in range(nk.value)], float(fm.value)

    def get_depth(self):
        out = np.empty(self.W * self.H, np.float32)
        _check(lib().piko_raster_get_depth(self._h, _p(out)), "get_depth")
        return out.reshape(self.H, self.W)

    def get_bins(self):
        """Raster-stage bin queues (Stage::getData, stage.h:331-333) as CSR."""
        nb = ((self.W + self.T - 1) // self.T) * ((self.H + self.T - 1) // self.T)
        start = np.zeros(nb + 1, np.int64)
        total = lib().piko_raster_get_bins(self._h, _p(start), None, 0)
        if total < 0:
            raise PikoError(int(total), "get_bins")
        tris = np.zeros(max(int(total), 1), np.int32)
        total = lib().piko_raster_get_bins(self._h, _p(start), _p(tris), int(total))
        return start, tris[: int(total)]

    def stats(self):
        st = Stats()
        _check(lib().piko_raster_get_stats(self._h, ctypes.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def destroy(self):
        if self._h is not None:
            lib().piko_raster_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


class BatchPipe:
    """Many screens (jobs) over one resident triangle buffer, each frame of the
    whole batch in one kernel sequence (piko_batch_* in include/piko_hip.h)."""

    def __init__(self, pos, attr, kind, subpixel_bits=8, max_jobs=1024, max_tris=0, max_pixels=1 << 20,
                 pair_capacity=0, on_device=False, raster_mode=0, part_pairs=0, waves_per_cu=0, queue_cover=0):
        self.kind = PIKO_INPUT_PRETRANSFORMED if kind == "pretransformed" else PIKO_INPUT_WORLD
        self.S = subpixel_bits
        if not on_device:
            pos = np.ascontiguousarray(pos, np.float32)
            attr = np.ascontiguousarray(attr)
            self._keep = (pos, attr)
            pp, ap, n = _p(pos), _p(attr), pos.shape[0]
        else:
            pp, ap, n = pos, attr, max_tris
        self.input_tris = int(n)
        opt = BatchOptions(self.kind, subpixel_bits, max_jobs, max(max_tris, 1), max_pixels, pair_capacity,
                           raster_mode, part_pairs, waves_per_cu, queue_cover)
        h = ctypes.c_void_p()
        _check(lib().piko_batch_create(ctypes.byref(h), pp, ap, int(n), 1 if on_device else 0, ctypes.byref(opt)),
               "batch_create")
        self._h = h
        self.jobs = []

    def submit(self, jobs):
        """jobs: list of (count, src_base, bin_size, ConstantState[, waves_per_cu, queue_depth[, queue_cover]])."""
        arr = (RasterJob * len(jobs))()
        for k, jb in enumerate(jobs):
            count, src, T, cs = jb[:4]
            arr[k].count, arr[k].src_base, arr[k].bin_size, arr[k].cs = int(count), int(src), int(T), cs
            if len(jb) > 4:
                arr[k].waves_per_cu = int(jb[4])
                arr[k].queue_depth = float(jb[5])
            if len(jb) > 6:
                arr[k].queue_cover = int(jb[6])
        self.jobs = [tuple(jb[:4]) for jb in jobs]
        _check(lib().piko_batch_submit(self._h, arr, len(jobs)), "batch_submit")

    def query(self):
        return lib().piko_batch_query(self._h) == 1

    def wait(self):
        _check(lib().piko_batch_wait(self._h), "batch_wait")

    def run_frames(self, n):
        """Re-runs the last submitted batch n times back to back."""
        _check(lib().piko_raster_run_frames(self._h, int(n)), "run_frames")

    def profile_frames(self, n, cap=64):
        return RasterPipe.profile_frames(self, n, cap)

    def color(self, j):
        cs = self.jobs[j][3]
        out = np.empty(cs.screenSizeX * cs.screenSizeY, np.uint32)
        _check(lib().piko_batch_get_color(self._h, int(j), _p(out)), "batch_get_color")
        return out.reshape(cs.screenSizeY, cs.screenSizeX)

    def depth(self, j):
        cs = self.jobs[j][3]
        out = np.empty(cs.screenSizeX * cs.screenSizeY, np.float32)
        _check(lib().piko_batch_get_depth(self._h, int(j), _p(out)), "batch_get_depth")
        return out.reshape(cs.screenSizeY, cs.screenSizeX)

    def stats(self):
        return RasterPipe.stats(self)

    def destroy(self):
        if self._h is not None:
            lib().piko_raster_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


class ReyesPipe:
    """Split -> Dice -> Shade -> PikoScreen on the GPU: the pipe pikoc generates
    for ReyesPipe (reyesPipe/reyesPipe.h:12-40), over the piko_reyes_* C-ABI.
    `scene` is a scenes.PatchScene (patches, camera and Reyes settings)."""

    def __init__(self):
        self._h = None

    def allocate(self, scene, roots_per_wave=0, keep_samples=False, max_splits=20, sampling=SAMPLE_AUTO, tile=0):
        """`sampling`: SAMPLE_GLOBAL (per-sample atomics in a visibility
        buffer), SAMPLE_TILED (per-tile LDS compositing) or SAMPLE_AUTO
        (global); `tile`: the tiled path's tile edge, 16 or 32 (0 = 16)."""
        if self._h is not None:
            raise RuntimeError("pipe already allocated")
        self.W, self.H, self.spp = scene.W, scene.H, scene.spp
        self.cs = ConstantState.for_scene(scene)
        opt = ReyesOptions(float(scene.threshold), int(max_splits), int(scene.dice), int(scene.spp),
                           float(scene.disp), int(roots_per_wave), 1 if keep_samples else 0, int(sampling), int(tile))
        cp = np.ascontiguousarray(scene.cp, np.float32)
        h = ctypes.c_void_p()
        _check(lib().piko_reyes_allocate(ctypes.byref(h), ctypes.byref(self.cs), _p(cp), int(cp.shape[0]),
                                         ctypes.byref(opt)), "reyes_allocate")
        self._h = h
        return self

    def allocate_device(self, scene, d_patches, count, roots_per_wave=0, keep_samples=False, max_splits=20,
                        sampling=SAMPLE_AUTO, tile=0):
        """As allocate, with the patches already in device memory
        (piko_reyes_allocate_device): d_patches is the device address of
        count patches of 16 control points x 4 floats; `scene` supplies the
        camera, screen and Reyes settings only."""
        if self._h is not None:
            raise RuntimeError("pipe already allocated")
        self.W, self.H, self.spp = scene.W, scene.H, scene.spp
        self.cs = ConstantState.for_scene(scene)
        opt = ReyesOptions(float(scene.threshold), int(max_splits), int(scene.dice), int(scene.spp),
                           float(scene.disp), int(roots_per_wave), 1 if keep_samples else 0, int(sampling), int(tile))
        h = ctypes.c_void_p()
        _check(lib().piko_reyes_allocate_device(ctypes.byref(h), ctypes.byref(self.cs), ctypes.c_void_p(d_patches),
                                                int(count), ctypes.byref(opt)), "reyes_allocate_device")
        self._h = h
        return self

    def prepare(self, scene=None):
        if scene is not None:
            self.cs = ConstantState.for_scene(scene)
        _check(lib().piko_reyes_prepare(self._h, ctypes.byref(self.cs)), "reyes_prepare")

    def run_single(self):
        _check(lib().piko_reyes_run_single(self._h), "reyes_run_single")

    def run_frames(self, n):
        _check(lib().piko_reyes_run_frames(self._h, int(n)), "reyes_run_frames")

    def profile_frames(self, n, cap=8):
        ms = (ctypes.c_float * cap)()
        names = (ctypes.c_char_p * cap)()
        nk = ctypes.c_int()
        fm = ctypes.c_float()
        _check(lib().piko_reyes_profile_frames(self._h, int(n), ctypes.cast(ms, ctypes.c_void_p),
                                               ctypes.cast(names, ctypes.c_void_p), cap, ctypes.byref(nk),
                                               ctypes.byref(fm)), "reyes_profile_frames")
        return [(names[k].decode(), float(ms[k])) for k in range(nk.value)], float(fm.value)

    def color(self):
        """FP32 RGBA [H, W, 4], row 0 at the bottom."""
        out = np.empty((self.H, self.W, 4), np.float32)
        _check(lib().piko_reyes_get_color(self._h, _p(out)), "reyes_get_color")
        return out

    def depth(self):
        out = np.empty((self.H, self.W), np.float32)
        _check(lib().piko_reyes_get_depth(self._h, _p(out)), "reyes_get_depth")
        return out

    def samples(self):
        """Per-sample depth bits [H, W, spp] (allocate with keep_samples=True)."""
        out = np.empty((self.H, self.W, self.spp), np.int32)
        _check(lib().piko_reyes_get_samples(self._h, _p(out)), "reyes_get_samples")
        return out

    def leaves(self):
        """Split leaves in emission order: (cp [L, 64] f32, nu [L], nv [L])."""
        n = lib().piko_reyes_get_leaves(self._h, None, None, None, 0)
        if n < 0:
            raise PikoError(int(n), "reyes_get_leaves")
        cp = np.empty((max(n, 1), 64), np.float32)
        nu = np.empty(max(n, 1), np.int32)
        nv = np.empty(max(n, 1), np.int32)
        n = lib().piko_reyes_get_leaves(self._h, _p(cp), _p(nu), _p(nv), int(n))
        if n < 0:
            raise PikoError(int(n), "reyes_get_leaves")
        return cp[:n], nu[:n], nv[:n]

    def stats(self):
        st = ReyesStats()
        _check(lib().piko_reyes_get_stats(self._h, ctypes.byref(st)), "reyes_stats")
        return {k: getattr(st, k) for k, _ in ReyesStats._fields_}

    def destroy(self):
        if self._h is not None:
            lib().piko_reyes_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


def render_reyes(scene, keep_samples=False):
    """One Reyes frame; returns (rgba [H,W,4] f32, depth [H,W] f32, stats)."""
    with ReyesPipe().allocate(scene, keep_samples=keep_samples) as pipe:
        pipe.prepare()
        pipe.run_single()
        return pipe.color(), pipe.depth(), pipe.stats()


class DeviceArray:
    """A device copy of a numpy array through the runtime shim (piko_rt_malloc
    + piko_rt_memcpy_htod); `ptr` is its device address."""

    def __init__(self, arr):
        L = lib()
        L.piko_rt_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.piko_rt_memcpy_htod.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.piko_rt_free.argtypes = [ctypes.c_void_p]
        a = np.ascontiguousarray(arr)
        self.nbytes = a.nbytes
        self._d = ctypes.c_void_p()
        _check(L.piko_rt_malloc(ctypes.byref(self._d), max(a.nbytes, 4)), "rt_malloc")
        if a.nbytes:
            _check(L.piko_rt_memcpy_htod(self._d, _p(a), a.nbytes), "rt_memcpy_htod")

    @property
    def ptr(self):
        return self._d.value

    def free(self):
        if self._d:
            lib().piko_rt_free(self._d)
            self._d = ctypes.c_void_p()


def frame_noalloc(cs, d_pos, d_attr, count, kind, subpixel_bits, d_color, d_depth):
    """piko_raster_frame_noalloc: one DIRECT-path frame of `count` device-resident
    triangles into caller-owned device colour and depth buffers, with no
    device allocation (per-frame buffers in the library's fixed scratch)."""
    L = lib()
    L.piko_raster_frame_noalloc.argtypes = [ctypes.POINTER(ConstantState), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p]
    k = PIKO_INPUT_PRETRANSFORMED if kind == "pretransformed" else PIKO_INPUT_WORLD
    return L.piko_raster_frame_noalloc(ctypes.byref(cs), ctypes.c_void_p(d_pos), ctypes.c_void_p(d_attr),
                                       int(count), k, int(subpixel_bits), ctypes.c_void_p(d_color),
                                       ctypes.c_void_p(d_depth))


def concat_scenes(scene_list):
    """Concatenates scenes of one kind into one input buffer; returns
    (pos, attr, [(count, src_base, scene)])."""
    kind = scene_list[0].kind
    pos = np.concatenate([s.pos for s in scene_list]) if scene_list else np.zeros((0, 9), np.float32)
    attr = np.concatenate([(s.col if kind == "pretransformed" else s.nrm) for s in scene_list])
    out, base = [], 0
    for s in scene_list:
        out.append((s.n, base, s))
        base += s.n
    return pos, attr, out


def render(scene, bin_size=16, subpixel_bits=8):
    """One frame; returns (color[H,W] uint32, depth[H,W] float32, pipe stats)."""
    with RasterPipe().allocate(scene, bin_size, subpixel_bits) as pipe:
        pipe.prepare()
        pipe.run_single()
        return pipe.pikoScreen.getData(), pipe.get_depth(), pipe.stats()
