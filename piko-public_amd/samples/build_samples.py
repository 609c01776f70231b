"""Builds the C++ samples: pikoc_hip --target=HIP on each pipe header, then
g++ on the host program against libpiko_hip.so and libpiko_scene.so, and
against the pipe's own lib<Pipe>_stages.so when pikoc_hip compiled its
stage code (the *_compiled builds force that for the library's samples;
raster_lit's bodies are not the library's, so it is always compiled).

  python piko-public_amd/samples/build_samples.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
sys.path.insert(0, PKG)

# (sample dir, pipe header, binary, defines)
COMPILED = {"PIKO_STAGE_CODE": "COMPILED"}
SAMPLES = [
    ("raster", "rasterPipe.h", "raster_sample", {}),
    ("raster", "rasterPipe.h", "raster_sample_world", {"RASTER_WORLD_INPUT": "1"}),
    ("reyes", "reyesPipe.h", "reyes_sample", {}),
    ("raster", "rasterPipe.h", "raster_sample_compiled", dict(COMPILED)),
    ("raster", "rasterPipe.h", "raster_sample_world_compiled", dict(COMPILED, RASTER_WORLD_INPUT="1")),
    ("reyes", "reyesPipe.h", "reyes_sample_compiled", dict(COMPILED)),
    ("raster_lit", "rasterPipe.h", "raster_lit_sample", {}),
]


def build_one(sub, header, binary, defines, log=None, root=HERE):
    """pikoc_hip on root/sub/header, then the host program root/sub/main.cpp
    into root/sub/binary."""
    from pikoc_hip import compile_pipe
    src = os.path.join(root, sub)
    gen = os.path.join(src, "gen_" + binary)
    inc = os.path.join(PKG, "include")
    be = compile_pipe(os.path.join(src, header), gen, [inc], defines, log=log)
    out = os.path.join(src, binary)
    libdir = os.path.join(PKG, "piko_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"), "-I", inc, "-I", gen, "-I", src]
    cmd += ["-D%s=%s" % kv for kv in defines.items()]
    cmd += [os.path.join(src, "main.cpp"), "-L", libdir, "-lpiko_hip", "-lpiko_scene",
            "-Wl,-rpath,$ORIGIN/" + os.path.relpath(libdir, src), "-Wl,-rpath-link,/opt/rocm/lib"]
    if be.kind == "compiled":
        cmd += ["-L", gen, "-l%s" % be.prefix, "-Wl,-rpath,$ORIGIN/gen_" + binary]
    cmd += ["-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build_all(log=None):
    return [build_one(*s, log=log) for s in SAMPLES]


if __name__ == "__main__":
    for p in build_all(log=print):
        print(p)
