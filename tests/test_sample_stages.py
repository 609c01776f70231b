"""The sample pipes' stage bodies are the algorithm the kernel library runs:
executed serially on the host (tests/stage_host/run_stages.cpp, built with
g++ per setting of the sample macros), the raster pipe's vertex and raster
stages and the Reyes pipe's split, dice and shade stages produce the
oracle's frames bit for bit -- colour words and depth bits, FP32 RGBA and
every sample's depth."""
import os
import subprocess

import numpy as np
import pytest

from piko_amd import ConstantState, scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(REPO, "tests", "stage_host", "run_stages.cpp")
SAMPLES = os.path.join(REPO, "piko-public_amd", "samples")


def build(tmp_path, name, sample, defines):
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-Wall", "-D__PIKOC_DEVICE__",
           "-I", os.path.join(REPO, "include"), "-I", os.path.join(REPO, "piko-public_amd", "include"),
           "-I", os.path.join(SAMPLES, sample)]
    cmd += ["-D%s=%s" % kv for kv in defines.items()]
    subprocess.run(cmd + [HARNESS, "-o", str(exe)], check=True)
    return exe


def run(exe, tmp_path, scene, arrays):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(np.array([scene.W, scene.H, scene.n], np.int32).tobytes())
        f.write(bytes(ConstantState.for_scene(scene)))
        for a in arrays:
            f.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([str(exe), str(src), str(dst)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return np.fromfile(dst, np.uint8), r.stdout


def aligned_grid(W=120, H=90, dx=7, dy=5):
    """Two triangles per dx x dy cell with every vertex on a pixel centre, so
    shared edges run through pixel centres (edge inclusion decides them) and
    every triangle has the same depth (the tie rule decides the colour)."""
    tris = []
    for y in range(0, H - dy, dy):
        for x in range(0, W - dx, dx):
            a, b = (x + 0.5, y + 0.5), (x + dx + 0.5, y + 0.5)
            c, d = (x + dx + 0.5, y + dy + 0.5), (x + 0.5, y + dy + 0.5)
            tris += [(a, b, c), (a, c, d)]
    pos = np.array([[v for p in t for v in (p[0], p[1], 0.5)] for t in tris], np.float32)
    col = (np.arange(3 * len(tris), dtype=np.uint32).reshape(-1, 3) * 2654435761) | np.uint32(0xff000000)
    return scenes.TriScene(pos, col.astype(np.uint32), W, H)


@pytest.mark.parametrize("case", ["uniform", "equal_depth", "aligned", "heightfield"])
def test_raster_stage_bodies_equal_oracle(oracle_mod, tmp_path, case):
    if case == "aligned":
        scene = aligned_grid()
        defines = {"GORAUD": "1", "PIKO_SUBPIXEL_BITS": "8", "VTX_PRETRANSFORM": "1"}
        arrays = (scene.pos, scene.col)
    elif case != "heightfield":
        # equal_depth: every triangle at one depth, so each pixel's colour is
        # decided by the tie rule (the later triangle wins)
        z = (0.1, 1.0) if case == "uniform" else (0.5, 0.5)
        scene = scenes.tris_uniform(4, 6000, 200, 150, z_lo=z[0], z_hi=z[1])
        defines = {"GORAUD": "1", "PIKO_SUBPIXEL_BITS": "8", "VTX_PRETRANSFORM": "1"}
        arrays = (scene.pos.astype(np.float32), scene.col.astype(np.uint32))
    else:
        scene = scenes.heightfield(seed=2, grid=60, W=240, H=136)
        defines = {"GORAUD": "1", "PIKO_SUBPIXEL_BITS": "8"}
        arrays = (scene.pos.astype(np.float32), scene.nrm.astype(np.float32))
    exe = build(tmp_path, "raster_" + case, "raster", defines)
    raw, _ = run(exe, tmp_path, scene, arrays)
    npix = scene.W * scene.H
    color = raw[:4 * npix].view(np.uint32).reshape(scene.H, scene.W)
    depth = raw[4 * npix:].view(np.int32).reshape(scene.H, scene.W)
    # bins of 8 px: the sample's RASTER_BINSIZE
    oc, od = oracle_mod.frame(scene, 8, 8)[:2]
    assert (oc != 0xff663313).sum() > npix // 4, "the scene covers too little to test"
    assert np.array_equal(color, oc), "colour differs at %d pixels" % int((color != oc).sum())
    assert np.array_equal(depth, od.view(np.int32)), "depth differs"


@pytest.mark.parametrize("dice,spp,disp", [(0, 1, 0.0), (8, 4, 0.05), (4, 1, 0.05)])
def test_reyes_stage_bodies_equal_oracle(oracle_mod, tmp_path, dice, spp, disp):
    scene = scenes.sphere_patches(seed=3, nlat=6, nlon=8, W=200, H=120, dice=dice, spp=spp, disp=disp,
                                  threshold=8.0)
    defines = {"RUN_REYES": "1", "THRESHOLD_PIXELS": "%rf" % scene.threshold, "PIKO_REYES_DICE": str(dice),
               "PIKO_REYES_SPP": str(spp), "PIKO_REYES_DISPLACEMENT": "%rf" % disp}
    exe = build(tmp_path, "reyes_%d_%d" % (dice, spp), "reyes", defines)
    raw, out = run(exe, tmp_path, scene, (scene.cp.astype(np.float32),))
    npix = scene.W * scene.H
    rgba = raw[:16 * npix].view(np.float32).reshape(scene.H, scene.W, 4)
    z = raw[16 * npix:].view(np.int32).reshape(scene.H, scene.W, spp)
    orgba, oz, nup = oracle_mod.reyes_frame(scene)
    assert ("%d micropolygons" % nup) in out
    assert (oz != np.float32(1.0).view(np.int32)).sum() > npix * spp // 10
    assert np.array_equal(z, oz), "sample depth differs at %d samples" % int((z != oz).sum())
    assert np.array_equal(rgba.view(np.int32), orgba.view(np.int32)), "colour differs"


def test_lit_sample_bodies_share_the_oracle_coverage_and_depth(oracle_mod, tmp_path):
    """The per-pixel-lit sample changes only the shading: run on the host,
    its bodies give the oracle's depth words bit for bit, and a colour at
    every covered pixel that its own normals decide (a Gouraud colour only
    where the two agree)."""
    scene = scenes.heightfield(seed=2, grid=60, W=240, H=136)
    exe = build(tmp_path, "raster_lit", "raster_lit", {"PIKO_SUBPIXEL_BITS": "8"})
    raw, out = run(exe, tmp_path, scene, (scene.pos.astype(np.float32), scene.nrm.astype(np.float32)))
    npix = scene.W * scene.H
    color = raw[:4 * npix].view(np.uint32).reshape(scene.H, scene.W)
    depth = raw[4 * npix:].view(np.int32).reshape(scene.H, scene.W)
    oc, od = oracle_mod.frame(scene, 8, 8)[:2]
    assert np.array_equal(depth, od.view(np.int32)), "depth differs"
    covered = oc != 0xff663313
    assert np.array_equal(color == 0xff663313, ~covered), "coverage differs"
    assert (color[covered] != oc[covered]).mean() > 0.5, "per-pixel shading should differ from Gouraud"
