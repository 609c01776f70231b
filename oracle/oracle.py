"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker for the HIP raster path: tests, the smoke entry
and the CPU-baseline leg of the benchmark import this module; the product
never does.  Functions mirror piko_oracle.h (each C function cites the
reference file:line it restates).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        stri = [ctypes.POINTER(_Stri)]
        L.oracle_vs_pretransformed.argtypes = [P, P, I, I, I, I] + stri
        L.oracle_vs_world.argtypes = [P, P, I, I, I, I, P, P, P, P] + stri
        L.oracle_bin.argtypes = stri + [I, I, I, I, I, P, P, LL]
        L.oracle_bin.restype = LL
        L.oracle_raster.argtypes = stri + [I, I, I, I, P, P, P, P]
        L.oracle_raster_bruteforce.argtypes = stri + [I, I, I, I, P, P]
        L.oracle_frame_pretransformed.argtypes = [P, P, I, I, I, I, I, I, P, P]
        L.oracle_frame_pretransformed.restype = LL
        L.oracle_frame_world.argtypes = [P, P, I, I, I, I, I, P, P, P, P, I, P, P]
        L.oracle_frame_world.restype = LL
        L.oracle_fragment_stats.argtypes = stri + [I, I, I, I, P, P]
        L.oracle_to_abgr.argtypes = [ctypes.c_float] * 3
        L.oracle_to_abgr.restype = ctypes.c_uint32
        L.oracle_from_abgr.argtypes = [ctypes.c_uint32, P]
        L.oracle_fixpt.argtypes = [ctypes.c_float, I]
        L.oracle_fixpt.restype = ctypes.c_int32
        L.oracle_reyes_split.argtypes = [P, I, I, I, P, ctypes.c_float, I, P, P, P, LL]
        L.oracle_reyes_split.restype = LL
        L.oracle_reyes_dice.argtypes = [P, P, P, LL, I, I, I, P, ctypes.c_float, P, LL]
        L.oracle_reyes_dice.restype = LL
        L.oracle_reyes_shade.argtypes = [P, LL, I, I, I, P, P]
        L.oracle_reyes_frame.argtypes = [P, I, I, I, P, ctypes.c_float, I, I, ctypes.c_float, P, P]
        L.oracle_reyes_frame.restype = LL
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Stri(ctypes.Structure):
    _fields_ = [("xy", ctypes.c_void_p), ("z", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("keep", ctypes.c_void_p)]


class Setup:
    """raster_stri in SoA form (rasterTypes.h:36-71)."""

    def __init__(self, n):
        self.n = n
        self.xy = np.zeros((n, 6), np.int32)
        self.z = np.zeros((n, 3), np.float32)
        self.col = np.zeros((n, 3), np.uint32)
        self.keep = np.zeros(n, np.uint8)

    def args(self):
        self._s = _Stri(self.xy.ctypes.data, self.z.ctypes.data, self.col.ctypes.data, self.keep.ctypes.data)
        return [ctypes.byref(self._s)]


def vertex_stage(scene, S):
    s = Setup(scene.n)
    if scene.kind == "pretransformed":
        pos = np.ascontiguousarray(scene.pos, np.float32)
        col = np.ascontiguousarray(scene.col, np.uint32)
        lib().oracle_vs_pretransformed(_p(pos), _p(col), scene.n, scene.W, scene.H, S, *s.args())
    else:
        pos = np.ascontiguousarray(scene.pos, np.float32)
        nrm = np.ascontiguousarray(scene.nrm, np.float32)
        vp = np.ascontiguousarray(scene.view_proj, np.float32)
        l, m, a = (np.ascontiguousarray(v, np.float32) for v in (scene.light, scene.matcol, scene.ambient))
        lib().oracle_vs_world(_p(pos), _p(nrm), scene.n, scene.W, scene.H, S, _p(vp), _p(l), _p(m), _p(a),
                              *s.args())
    return s


def bins(setup, W, H, S, T):
    nb = ((W + T - 1) // T) * ((H + T - 1) // T)
    start = np.zeros(nb + 1, np.int64)
    lib().oracle_bin(*setup.args(), setup.n, W, H, S, T, _p(start), None, 0)
    tris = np.zeros(max(int(start[-1]), 1), np.int32)
    total = lib().oracle_bin(*setup.args(), setup.n, W, H, S, T, _p(start), _p(tris), int(start[-1]))
    assert total == start[-1]
    return start, tris[: int(total)]


def raster(setup, W, H, S, T, start, tris):
    color = np.zeros(W * H, np.uint32)
    depth = np.zeros(W * H, np.float32)
    tris = np.ascontiguousarray(tris if len(tris) else np.zeros(1, np.int32), np.int32)
    lib().oracle_raster(*setup.args(), W, H, S, T, _p(start), _p(tris), _p(color), _p(depth))
    return color.reshape(H, W), depth.reshape(H, W)


def raster_bruteforce(setup, W, H, S):
    color = np.zeros(W * H, np.uint32)
    depth = np.zeros(W * H, np.float32)
    lib().oracle_raster_bruteforce(*setup.args(), setup.n, W, H, S, _p(color), _p(depth))
    return color.reshape(H, W), depth.reshape(H, W)


def frame(scene, S, T, threads=1):
    """Whole frame (vertex stage, bin walk, raster) -> (color[H,W], depth[H,W], pairs)."""
    W, H = scene.W, scene.H
    color = np.zeros(W * H, np.uint32)
    depth = np.zeros(W * H, np.float32)
    if scene.kind == "pretransformed":
        pos = np.ascontiguousarray(scene.pos, np.float32)
        col = np.ascontiguousarray(scene.col, np.uint32)
        pairs = lib().oracle_frame_pretransformed(_p(pos), _p(col), scene.n, W, H, S, T, threads,
                                                  _p(color), _p(depth))
    else:
        pos = np.ascontiguousarray(scene.pos, np.float32)
        nrm = np.ascontiguousarray(scene.nrm, np.float32)
        vp = np.ascontiguousarray(scene.view_proj, np.float32)
        l, m, a = (np.ascontiguousarray(v, np.float32) for v in (scene.light, scene.matcol, scene.ambient))
        pairs = lib().oracle_frame_world(_p(pos), _p(nrm), scene.n, W, H, S, T, _p(vp), _p(l), _p(m), _p(a),
                                         threads, _p(color), _p(depth))
    return color.reshape(H, W), depth.reshape(H, W), pairs


def fragment_stats(setup, W, H, S):
    f = ctypes.c_longlong()
    c = ctypes.c_longlong()
    lib().oracle_fragment_stats(*setup.args(), setup.n, W, H, S, ctypes.byref(f), ctypes.byref(c))
    return f.value, c.value


def to_abgr(r, g, b):
    return int(lib().oracle_to_abgr(r, g, b))


def from_abgr(c):
    out = np.zeros(3, np.float32)
    lib().oracle_from_abgr(c, _p(out))
    return out


def fixpt(v, S):
    return int(lib().oracle_fixpt(v, S))


# ---- Reyes ---------------------------------------------------------------

UPOLY = np.dtype([("x", np.float32, 4), ("y", np.float32, 4), ("z", np.float32), ("rgb", np.float32, 3)])


def reyes_split(cp, W, H, vp, threshold, max_splits=20):
    """Leaves in (level, position) order: (leaves [L,64] f32, nu [L], nv [L])."""
    cp = np.ascontiguousarray(cp, np.float32)
    vp = np.ascontiguousarray(vp, np.float32)
    L = lib()
    n = L.oracle_reyes_split(_p(cp), cp.shape[0], W, H, _p(vp), threshold, max_splits, None, None, None, 0)
    leaves = np.empty((max(n, 1), 64), np.float32)
    nu = np.empty(max(n, 1), np.int32)
    nv = np.empty(max(n, 1), np.int32)
    L.oracle_reyes_split(_p(cp), cp.shape[0], W, H, _p(vp), threshold, max_splits, _p(leaves), _p(nu), _p(nv), n)
    return leaves[:n], nu[:n], nv[:n]


def reyes_dice(leaves, nu, nv, dice, W, H, vp, disp):
    leaves = np.ascontiguousarray(leaves, np.float32)
    nu = np.ascontiguousarray(nu, np.int32)
    nv = np.ascontiguousarray(nv, np.int32)
    vp = np.ascontiguousarray(vp, np.float32)
    L = lib()
    n = L.oracle_reyes_dice(_p(leaves), _p(nu), _p(nv), leaves.shape[0], dice, W, H, _p(vp), disp, None, 0)
    out = np.empty(max(n, 1), UPOLY)
    L.oracle_reyes_dice(_p(leaves), _p(nu), _p(nv), leaves.shape[0], dice, W, H, _p(vp), disp, _p(out), n)
    return out[:n]


def reyes_shade(upolys, W, H, spp):
    """(rgba [H,W,4] f32, sample z bits [H,W,spp] i32)."""
    upolys = np.ascontiguousarray(upolys, UPOLY)
    rgba = np.empty((H, W, 4), np.float32)
    z = np.empty((H, W, spp), np.int32)
    lib().oracle_reyes_shade(_p(upolys), upolys.shape[0], W, H, spp, _p(rgba), _p(z))
    return rgba, z


def reyes_frame(scene):
    """Whole Reyes frame of a PatchScene: (rgba [H,W,4], sample z bits, micropolygons)."""
    cp = np.ascontiguousarray(scene.cp, np.float32)
    vp = np.ascontiguousarray(scene.view_proj, np.float32)
    rgba = np.empty((scene.H, scene.W, 4), np.float32)
    z = np.empty((scene.H, scene.W, scene.spp), np.int32)
    nup = lib().oracle_reyes_frame(_p(cp), cp.shape[0], scene.W, scene.H, _p(vp), scene.threshold, scene.dice,
                                   scene.spp, scene.disp, _p(rgba), _p(z))
    return rgba, z, nup
