"""Python view of the scene generators (libpiko_scene.so)."""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpiko_scene.so")
        if not os.path.exists(path):
            raise RuntimeError("libpiko_scene.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.piko_scene_tris_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_float, P, P]
        L.piko_scene_heightfield.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_int, P, P]
        L.piko_scene_heightfield_freq.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                  ctypes.c_float, P, P]
        L.piko_scene_heightfield_count.restype = ctypes.c_longlong
        L.piko_camera_view_proj.argtypes = [P, P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_float, P]
        L.piko_scene_soup_world.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_float, P, P]
        L.piko_scene_batch_spec.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, P, P, P]
        L.piko_scene_sphere_patches.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_float, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class TriScene:
    """Pretransformed screen-space soup: pos [n,9] f32, col [n,3] u32."""

    kind = "pretransformed"

    def __init__(self, pos, col, W, H, seed=None):
        self.pos, self.col, self.W, self.H, self.seed = pos, col, W, H, seed

    @property
    def n(self):
        return self.pos.shape[0]


class WorldScene:
    """World-space soup with normals and a camera: pos/nrm [n,9] f32."""

    kind = "world"

    def __init__(self, pos, nrm, W, H, view_proj, light, matcol, ambient, seed=None):
        self.pos, self.nrm, self.W, self.H = pos, nrm, W, H
        self.view_proj = view_proj
        self.light, self.matcol, self.ambient, self.seed = light, matcol, ambient, seed

    @property
    def n(self):
        return self.pos.shape[0]


def tris_uniform(seed, n, W, H, area_median=20.0, sigma=1.0, z_lo=0.1, z_hi=1.0):
    pos = np.empty((n, 9), np.float32)
    col = np.empty((n, 3), np.uint32)
    if lib().piko_scene_tris_uniform(seed, n, W, H, area_median, sigma, z_lo, z_hi, _p(pos), _p(col)) != 0:
        raise ValueError("bad scene arguments")
    return TriScene(pos, col, W, H, seed)


# Light and material of the Gouraud vertex lighting (the constants of
# rasterPipelineFixPt/main.cpp:317-328 and vertexShader.pikostage:94-95).
DEFAULT_LIGHT = (0.5773, 0.5773, 0.5773)
DEFAULT_MATCOL = (0.9, 0.9, 0.6)
DEFAULT_AMBIENT = (0.15, 0.15, 0.30)


def view_proj(eye, target, up, fovy_deg, aspect, znear, zfar):
    m = np.empty(16, np.float32)
    e, t, u = (np.asarray(v, np.float32) for v in (eye, target, up))
    if lib().piko_camera_view_proj(_p(e), _p(t), _p(u), fovy_deg, aspect, znear, zfar, _p(m)) != 0:
        raise ValueError("bad camera")
    return m


# Height-field camera pose and terrain: a low view across a rugged terrain
# (first noise octave at 16 lattice cells per unit) so hills overlap.  Over
# the 2M-triangle grid at 1080p this gives a mean depth complexity of 3.94
# (8.16M fragments over 2.07M covered pixels, 99.8% of the screen); a sweep
# of poses and lattice frequencies is recorded with the benchmark profiles.
HF_EYE = (0.0, 0.12, 0.7)
HF_TARGET = (0.0, -0.3, 0.0)
HF_FREQ = 16.0
HF_UP = (0.0, 1.0, 0.0)


def heightfield(seed=2, grid=1000, W=1920, H=1080, amplitude=0.2, octaves=6,
                eye=HF_EYE, target=HF_TARGET, fovy=60.0, znear=0.02, zfar=4.0, base_freq=HF_FREQ):
    n = int(lib().piko_scene_heightfield_count(grid))
    pos = np.empty((n, 9), np.float32)
    nrm = np.empty((n, 9), np.float32)
    if lib().piko_scene_heightfield_freq(seed, grid, amplitude, octaves, base_freq, _p(pos), _p(nrm)) != 0:
        raise ValueError("bad heightfield arguments")
    vp = view_proj(eye, target, HF_UP, fovy, W / H, znear, zfar)
    return WorldScene(pos, nrm, W, H, vp, np.asarray(DEFAULT_LIGHT, np.float32),
                      np.asarray(DEFAULT_MATCOL, np.float32), np.asarray(DEFAULT_AMBIENT, np.float32), seed)


def soup_world(seed, n, extent=1.0, size_median=0.01, sigma=1.0):
    """World-space soup: pos/nrm [n, 9]."""
    pos = np.empty((n, 9), np.float32)
    nrm = np.empty((n, 9), np.float32)
    if lib().piko_scene_soup_world(seed, n, extent, size_median, sigma, _p(pos), _p(nrm)) != 0:
        raise ValueError("bad soup arguments")
    return pos, nrm


def batch_spec(base_seed=1000, nscenes=256, n_lo=1000, n_hi=500000, px_lo_edge=256, px_hi_edge=2048):
    n = np.empty(nscenes, np.int32)
    w = np.empty(nscenes, np.int32)
    h = np.empty(nscenes, np.int32)
    if lib().piko_scene_batch_spec(base_seed, nscenes, n_lo, n_hi, px_lo_edge, px_hi_edge,
                                   _p(n), _p(w), _p(h)) != 0:
        raise ValueError("bad batch arguments")
    return [(base_seed + s, int(n[s]), int(w[s]), int(h[s])) for s in range(nscenes)]


class PatchScene:
    """Reyes input: cp [P, 64] f32 (16 control points x xyzw per patch, the
    piko_patch CP layout) with a camera and the Reyes settings: split
    threshold in pixels, dice grid edge (0 = the reference's nu/nv), samples
    per pixel (1 or 4) and displacement amplitude."""

    kind = "patches"

    def __init__(self, cp, W, H, view_proj, threshold=16.0, dice=8, spp=4, disp=0.05, seed=None):
        self.cp, self.W, self.H, self.view_proj = cp, W, H, view_proj
        self.threshold, self.dice, self.spp, self.disp, self.seed = threshold, dice, spp, disp, seed

    @property
    def n(self):
        return self.cp.shape[0]


# Reyes sphere camera: the unit sphere seen from distance 3 (about 700 px
# across at 720p), depth range [1, 5] so every displaced point is in front.
SPHERE_EYE = (0.4, 0.9, 2.8)
SPHERE_TARGET = (0.0, 0.0, 0.0)


def sphere_patches(seed=3, nlat=64, nlon=64, noise=0.05, W=1280, H=720, eye=SPHERE_EYE, fovy=40.0,
                   threshold=16.0, dice=8, spp=4, disp=0.05):
    cp = np.empty((nlat * nlon, 64), np.float32)
    if lib().piko_scene_sphere_patches(seed, nlat, nlon, noise, _p(cp)) != 0:
        raise ValueError("bad patch arguments")
    vp = view_proj(eye, SPHERE_TARGET, (0.0, 1.0, 0.0), fovy, W / H, 1.0, 5.0)
    return PatchScene(cp, W, H, vp, threshold, dice, spp, disp, seed)


def algorithmic_bytes(scene):
    """Compulsory HBM bytes of one frame: input read once + the framebuffer
    written once (raster: colour + depth, 4 B each per pixel; Reyes: FP32
    RGBA + FP32 depth per pixel)."""
    if scene.kind == "patches":
        return 256 * scene.n + 20 * scene.W * scene.H
    b_in = 48 if scene.kind == "pretransformed" else 72
    return b_in * scene.n + 8 * scene.W * scene.H


__all__ = ["TriScene", "WorldScene", "PatchScene", "tris_uniform", "heightfield", "sphere_patches", "view_proj",
           "batch_spec",
           "algorithmic_bytes", "DEFAULT_LIGHT", "DEFAULT_MATCOL", "DEFAULT_AMBIENT", "math"]
