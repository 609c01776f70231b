"""GPU: the C++ samples, compiled by pikoc_hip --target=HIP and g++ against
the C-ABI, render the same frames as the oracle, bit for bit (PikoScreen
words and depth), and one profiled frame launches exactly the kernels the
generated kernel map names.  Each sample runs both ways pikoc_hip binds a
pipe: on the library's kernels, and as its own stage bodies compiled for
gfx950 (the *_compiled builds).  The per-pixel-lit sample, whose bodies the
library does not implement, is checked against the same bodies run serially
on the host (tests/stage_host/run_stages.cpp)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from piko_amd import scenes

pytestmark = pytest.mark.gpu

SAMPLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "piko-public_amd", "samples")


def kernel_map(binary):
    """(kernels launched every frame, kernels launched only when needed) from
    the .kernels map pikoc_hip generated for the sample."""
    d, name = os.path.split(binary)
    path = os.path.join(SAMPLES, d, "gen_" + name, "__pikoCompiledPipe.kernels")
    always, maybe = [], []
    for line in open(path).read().splitlines()[1:]:
        ks = [t for t in line.split() if t.startswith("k_")]
        (maybe if line.rstrip().endswith("?") else always).extend(ks)
    return always, maybe


def run_sample(binary, args, tmp_path):
    path = os.path.join(SAMPLES, binary)
    if not os.path.exists(path):
        pytest.fail("sample %s is not built (run __graft_entry__.build())" % binary)
    print("sample", binary, args, flush=True)
    r = subprocess.run([path, "--runs", "3", "--dump", str(tmp_path)] + args, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    sys.stdout.flush()
    # the kernels one frame launched are the ones the generated map names
    launched = [k for entry in json.loads(r.stdout.strip().splitlines()[-1])["kernels"] for k in entry.split("+")]
    always, maybe = kernel_map(binary)
    assert set(always) <= set(launched) <= set(always + maybe), (launched, always, maybe)
    return np.fromfile(tmp_path / "color.bin", np.uint32), np.fromfile(tmp_path / "depth.bin", np.float32)


@pytest.mark.parametrize("binary", ["raster/raster_sample", "raster/raster_sample_compiled"])
def test_raster_sample_uniform(oracle_mod, tmp_path, binary):
    color, depth = run_sample(binary, [], tmp_path)
    sc = scenes.tris_uniform(1, 50000, 512, 512)
    oc, od = oracle_mod.frame(sc, 8, 8)[:2]
    assert np.array_equal(color, oc.reshape(-1))
    assert np.array_equal(depth.view(np.int32), od.reshape(-1).view(np.int32))


@pytest.mark.parametrize("binary", ["raster/raster_sample_world", "raster/raster_sample_world_compiled"])
def test_raster_sample_heightfield(oracle_mod, tmp_path, binary):
    color, depth = run_sample(binary, ["--scene", "heightfield"], tmp_path)
    sc = scenes.heightfield()
    oc, od = oracle_mod.frame(sc, 8, 8)[:2]
    assert np.array_equal(color, oc.reshape(-1))
    assert np.array_equal(depth.view(np.int32), od.reshape(-1).view(np.int32))


@pytest.mark.parametrize("binary", ["reyes/reyes_sample", "reyes/reyes_sample_compiled"])
def test_reyes_sample(oracle_mod, tmp_path, binary):
    color, depth = run_sample(binary, [], tmp_path)
    sc = scenes.sphere_patches()
    rgba, z, _ = oracle_mod.reyes_frame(sc)
    c = (rgba * np.float32(255.0)).astype(np.uint32)  # cvec2uintcolor, truncating
    words = (c[..., 3] << 24) | (c[..., 2] << 16) | (c[..., 1] << 8) | c[..., 0]
    assert np.array_equal(color, words.reshape(-1))
    assert np.array_equal(depth.view(np.int32), z.min(axis=2).reshape(-1))


def test_lit_sample_equals_its_bodies_run_on_the_host(oracle_mod, tmp_path):
    """The compiled per-pixel-lit stages on the 2M-triangle height field: the
    GPU frame equals the same stage bodies executed serially on the host,
    bit for bit, and its depth is the oracle's."""
    from test_sample_stages import build as build_host, run as run_host
    (tmp_path / "gpu").mkdir()
    color, depth = run_sample("raster_lit/raster_lit_sample", ["--scene", "heightfield"], tmp_path / "gpu")
    sc = scenes.heightfield()
    exe = build_host(tmp_path, "raster_lit_host", "raster_lit", {"PIKO_SUBPIXEL_BITS": "8"})
    raw, _ = run_host(exe, tmp_path, sc, (sc.pos.astype(np.float32), sc.nrm.astype(np.float32)))
    npix = sc.W * sc.H
    hc, hd = raw[:4 * npix].view(np.uint32), raw[4 * npix:].view(np.int32)
    assert np.array_equal(color, hc), "colour differs at %d pixels" % int((color != hc).sum())
    assert np.array_equal(depth.view(np.int32), hd)
    od = oracle_mod.frame(sc, 8, 8)[1]
    assert np.array_equal(depth.view(np.int32), od.reshape(-1).view(np.int32))
