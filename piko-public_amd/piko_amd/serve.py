"""Serving over resident scenes (piko_server_* in include/piko_hip.h).

Scenes are uploaded once; a trace of render requests (scene, camera,
resolution, arrival time) is replayed in real time by the native serving
loop, which batches whatever has arrived onto free batch pipes (HIP streams)
under a per-batch device-memory budget.
"""
import ctypes
import math

import numpy as np

from . import ConstantState, PikoError, _check, _p, lib, scenes

RESOLUTIONS = ((960, 540), (1280, 720), (1920, 1080))


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int), ("bin_size", ctypes.c_int), ("cs", ConstantState),
                ("arrival_s", ctypes.c_double)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("input_kind", ctypes.c_int), ("subpixel_bits", ctypes.c_int), ("pipes", ctypes.c_int),
                ("max_jobs", ctypes.c_int), ("batch_budget_bytes", ctypes.c_longlong)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_longlong), ("batches", ctypes.c_longlong), ("wall_s", ctypes.c_double),
                ("mean_batch", ctypes.c_double)]


def _bind():
    L = lib()
    if not getattr(L, "_serve_bound", False):
        P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        L.piko_server_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), P, P, LL, I, P, P, I,
                                         ctypes.POINTER(ServerOptions)]
        L.piko_server_replay.argtypes = [P, ctypes.POINTER(Request), I, P, P, ctypes.POINTER(ServerStats)]
        L.piko_server_destroy.argtypes = [P]
        L._serve_bound = True
    return L


def resident_scenes(nscenes=64, tris_per_scene=200_000, seed=7):
    """A mix of the two generators in world space: even scenes are triangle
    soups (uniform centroids in the unit cube), odd ones height fields.
    Returns (pos, nrm, scene_base, scene_count)."""
    pos_l, nrm_l, base, count = [], [], [], []
    off = 0
    for k in range(nscenes):
        if k % 2 == 0:
            p, n = scenes.soup_world(seed * 1000 + k, tris_per_scene)
        else:
            g = max(2, int(math.sqrt(tris_per_scene / 2)))
            hf = scenes.heightfield(seed=seed * 1000 + k, grid=g, W=64, H=64)
            p, n = hf.pos, hf.nrm
        pos_l.append(p)
        nrm_l.append(n)
        base.append(off)
        count.append(p.shape[0])
        off += p.shape[0]
    return (np.concatenate(pos_l), np.concatenate(nrm_l), np.asarray(base, np.int64),
            np.asarray(count, np.int32))


def poisson_trace(n, rate_hz, nscenes, seed=11, bin_size=8, resolutions=RESOLUTIONS):
    """n requests with exponential inter-arrival times (rate_hz), a uniform
    scene, a camera on a sphere of radius U[2,5] around the origin looking at
    it, vertical FOV U[40,90] deg and a resolution drawn from `resolutions`."""
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    reqs = (Request * n)()
    meta = []
    for i in range(n):
        sc = int(rng.integers(nscenes))
        rad = rng.uniform(2.0, 5.0)
        z = rng.uniform(-1.0, 1.0)
        ph = rng.uniform(0.0, 2 * math.pi)
        r = math.sqrt(max(0.0, 1 - z * z))
        eye = (rad * r * math.cos(ph), rad * z, rad * r * math.sin(ph))
        up = (0.0, 1.0, 0.0) if abs(z) < 0.95 else (1.0, 0.0, 0.0)
        fov = rng.uniform(40.0, 90.0)
        W, H = resolutions[int(rng.integers(len(resolutions)))]
        vp = scenes.view_proj(eye, (0.0, 0.0, 0.0), up, fov, W / H, 0.05, 10.0)
        view = scenes.WorldScene(None, None, W, H, vp, np.asarray(scenes.DEFAULT_LIGHT, np.float32),
                                 np.asarray(scenes.DEFAULT_MATCOL, np.float32),
                                 np.asarray(scenes.DEFAULT_AMBIENT, np.float32))
        reqs[i].scene = sc
        reqs[i].bin_size = bin_size
        reqs[i].cs = ConstantState.for_scene(view)
        reqs[i].arrival_s = float(arr[i])
        meta.append((sc, view))
    return reqs, meta


class Server:
    def __init__(self, pos, nrm, scene_base, scene_count, pipes=4, max_jobs=64, budget_bytes=8 << 30,
                 subpixel_bits=8):
        L = _bind()
        self.pos = np.ascontiguousarray(pos, np.float32)
        self.nrm = np.ascontiguousarray(nrm, np.float32)
        self.base = np.ascontiguousarray(scene_base, np.int64)
        self.count = np.ascontiguousarray(scene_count, np.int32)
        opt = ServerOptions(1, subpixel_bits, pipes, max_jobs, budget_bytes)
        h = ctypes.c_void_p()
        _check(L.piko_server_create(ctypes.byref(h), _p(self.pos), _p(self.nrm), int(self.pos.shape[0]), 0,
                                    _p(self.base), _p(self.count), len(self.base), ctypes.byref(opt)),
               "server_create")
        self._h = h

    def replay(self, reqs, keep=False):
        """Returns (done_s [n], stats dict, colours or None).  keep: True for
        every request's framebuffer, or a list of request indices."""
        n = len(reqs)
        done = np.zeros(n, np.float64)
        cols = None
        keep_arr = None
        if keep is not False:
            idx = set(range(n)) if keep is True else set(int(k) for k in keep)
            cols = [np.empty(r.cs.screenSizeX * r.cs.screenSizeY, np.uint32) if i in idx else None
                    for i, r in enumerate(reqs)]
            keep_arr = (ctypes.c_void_p * n)(*[c.ctypes.data if c is not None else None for c in cols])
        st = ServerStats()
        rc = _bind().piko_server_replay(self._h, reqs, n, _p(done), keep_arr, ctypes.byref(st))
        if rc != 0:
            raise PikoError(rc, "server_replay")
        arr = np.array([r.arrival_s for r in reqs])
        lat = done - arr
        stats = {"requests": st.requests, "batches": st.batches, "wall_s": st.wall_s,
                 "mean_batch": st.mean_batch,
                 "frames_per_s": n / st.wall_s if st.wall_s > 0 else 0.0,
                 "p50_ms": float(np.percentile(lat, 50) * 1e3) if n else 0.0,
                 "p99_ms": float(np.percentile(lat, 99) * 1e3) if n else 0.0}
        if keep is not False:
            cols = [c.reshape(r.cs.screenSizeY, r.cs.screenSizeX) if c is not None else None
                    for c, r in zip(cols, reqs)]
        return done, stats, cols

    def destroy(self):
        if self._h is not None:
            _bind().piko_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()
