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


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("done_s", ctypes.c_double), ("status", ctypes.c_int)]


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
        L.piko_server_submit.argtypes = [P, ctypes.POINTER(Request), P, ctypes.POINTER(LL)]
        L.piko_server_poll.argtypes = [P, ctypes.POINTER(Completion), I, I]
        L._serve_bound = True
    return L


def soup_size(tris):
    """Median circumradius of a soup scene's triangles: scaled as 1/sqrt(N) so
    that the projected depth complexity stays near config 1's (50,000
    triangles of median 20 px^2 on 512^2, about 6.3) at any triangle count --
    6.3 to 8.3 for 8.7M triangles at 1080p from 2 to 5 units (oracle coverage
    counts)."""
    return 0.007 * math.sqrt(300_000 / max(tris, 1))


def scene_geometry(k, tris, seed=7):
    """Scene k of the resident mix in world space: even k a triangle soup
    (uniform centroids in [-1, 1]^3, triangle size from soup_size), odd k a
    height field of about `tris` triangles.  Returns (pos, nrm), [n, 9]
    float32 each."""
    if k % 2 == 0:
        return scenes.soup_world(seed * 1000 + k, tris, 1.0, soup_size(tris))
    g = max(2, int(math.sqrt(tris / 2)))
    hf = scenes.heightfield(seed=seed * 1000 + k, grid=g, W=64, H=64)
    return hf.pos, hf.nrm


def scene_count(k, tris):
    return tris if k % 2 == 0 else 2 * max(2, int(math.sqrt(tris / 2))) ** 2


def resident_scenes(nscenes=64, tris_per_scene=200_000, seed=7, sizes=None):
    """The mix of the two generators on the host (scene k has sizes[k]
    triangles, or tris_per_scene).  Returns (pos, nrm, scene_base,
    scene_count)."""
    pos_l, nrm_l, base, count = [], [], [], []
    off = 0
    for k in range(nscenes):
        p, n = scene_geometry(k, sizes[k] if sizes else tris_per_scene, seed)
        pos_l.append(p)
        nrm_l.append(n)
        base.append(off)
        count.append(p.shape[0])
        off += p.shape[0]
    return (np.concatenate(pos_l), np.concatenate(nrm_l), np.asarray(base, np.int64),
            np.asarray(count, np.int32))


class DeviceScenes:
    """The same mix built scene by scene straight into device memory, so the
    host never holds more than one scene (the generators are counter-based:
    host_scene(k) regenerates scene k for checking)."""

    def __init__(self, nscenes=64, tris_per_scene=200_000, seed=7, sizes=None, log=None):
        L = lib()
        P, SZ = ctypes.c_void_p, ctypes.c_size_t
        L.piko_rt_malloc.argtypes = [ctypes.POINTER(P), SZ]
        L.piko_rt_memcpy_htod.argtypes = [P, P, SZ]
        L.piko_rt_free.argtypes = [P]
        self.seed = seed
        self.sizes = [sizes[k] if sizes else tris_per_scene for k in range(nscenes)]
        self.count = np.asarray([scene_count(k, t) for k, t in enumerate(self.sizes)], np.int32)
        self.base = np.concatenate([[0], np.cumsum(self.count.astype(np.int64))[:-1]]).astype(np.int64)
        self.tris = int(self.count.astype(np.int64).sum())
        self.nbytes = 72 * self.tris
        self._pos, self._nrm = P(), P()
        for ptr in (self._pos, self._nrm):
            _check(L.piko_rt_malloc(ctypes.byref(ptr), max(36 * self.tris, 4)), "rt_malloc")
        for k in range(nscenes):
            p, n = scene_geometry(k, self.sizes[k], seed)
            assert p.shape[0] == self.count[k]
            off = 36 * int(self.base[k])
            _check(L.piko_rt_memcpy_htod(P(self._pos.value + off), _p(p), p.nbytes), "rt_memcpy_htod")
            _check(L.piko_rt_memcpy_htod(P(self._nrm.value + off), _p(n), n.nbytes), "rt_memcpy_htod")
            if log and (k % 16 == 15 or k == nscenes - 1):
                log("resident scenes: %d of %d uploaded (%.1f GB)" % (
                    k + 1, nscenes, 72 * float(self.base[k] + self.count[k]) / 1e9))

    @property
    def pos_ptr(self):
        return self._pos.value

    @property
    def nrm_ptr(self):
        return self._nrm.value

    def host_scene(self, k):
        return scene_geometry(k, self.sizes[k], self.seed)

    def free(self):
        for ptr in (self._pos, self._nrm):
            if ptr:
                lib().piko_rt_free(ptr)
        self._pos, self._nrm = ctypes.c_void_p(), ctypes.c_void_p()


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
    """pos/nrm: host arrays of the scenes (uploaded), or a DeviceScenes as
    `pos` (adopted; nrm, scene_base and scene_count are then taken from it)."""

    def __init__(self, pos, nrm=None, scene_base=None, scene_count=None, pipes=4, max_jobs=64,
                 budget_bytes=8 << 30, subpixel_bits=8):
        L = _bind()
        opt = ServerOptions(1, subpixel_bits, pipes, max_jobs, budget_bytes)
        h = ctypes.c_void_p()
        if isinstance(pos, DeviceScenes):
            self.base = np.ascontiguousarray(pos.base, np.int64)
            self.count = np.ascontiguousarray(pos.count, np.int32)
            rc = L.piko_server_create(ctypes.byref(h), ctypes.c_void_p(pos.pos_ptr), ctypes.c_void_p(pos.nrm_ptr),
                                      pos.tris, 1, _p(self.base), _p(self.count), len(self.base), ctypes.byref(opt))
        else:
            self.pos = np.ascontiguousarray(pos, np.float32)
            self.nrm = np.ascontiguousarray(nrm, np.float32)
            self.base = np.ascontiguousarray(scene_base, np.int64)
            self.count = np.ascontiguousarray(scene_count, np.int32)
            rc = L.piko_server_create(ctypes.byref(h), _p(self.pos), _p(self.nrm), int(self.pos.shape[0]), 0,
                                      _p(self.base), _p(self.count), len(self.base), ctypes.byref(opt))
        _check(rc, "server_create")
        self._h = h
        self._live = {}

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

    def submit(self, req, keep=True):
        """Online submission (piko_server_submit): returns (id, framebuffer or
        None); the framebuffer is filled when the request completes."""
        col = np.empty(req.cs.screenSizeX * req.cs.screenSizeY, np.uint32) if keep else None
        rid = ctypes.c_longlong()
        one = Request()
        ctypes.pointer(one)[0] = req
        rc = _bind().piko_server_submit(self._h, ctypes.byref(one), _p(col) if col is not None else None,
                                        ctypes.byref(rid))
        if rc != 0:
            raise PikoError(rc, "server_submit")
        self._live[rid.value] = col  # keeps the buffer alive until its completion
        return rid.value, (col.reshape(req.cs.screenSizeY, req.cs.screenSizeX) if col is not None else None)

    def poll(self, cap=64, wait_ms=10):
        """Completions so far (piko_server_poll): [(id, done_s, status)]."""
        buf = (Completion * cap)()
        n = _bind().piko_server_poll(self._h, buf, cap, int(wait_ms))
        if n < 0:
            raise PikoError(n, "server_poll")
        out = [(buf[i].id, buf[i].done_s, buf[i].status) for i in range(n)]
        for rid, _, _ in out:
            self._live.pop(rid, None)
        return out

    def destroy(self):
        if self._h is not None:
            _bind().piko_server_destroy(self._h)
            self._h = None
            self._live = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()
