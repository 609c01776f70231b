"""Builds the native libraries of the MI355X Piko backend in-tree.

  piko_amd/libpiko_scene.so  scene generators (g++, host only)
  piko_amd/libpiko_hip.so    HIP kernels for gfx950 + the C-ABI runtime
  piko_amd/vertex_module.hsaco  a gfx950 code object (vertex stage + clear)
                              that the runtime shim's module calls load

Every call recompiles (hipcc cross-compiles gfx950 without a GPU).  Flags:
  -ffp-contract=off           no FMA contraction: the raster path replays the
                              reference CPU backend's unfused float ops
  -packed-fp32-ops disabled   no packed FP32 instructions are generated
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "piko_amd")
REPO = os.path.dirname(HERE)
INC = os.path.join(REPO, "include")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def _atomic_out(path):
    return path + ".tmp%d" % os.getpid()


def build_scene():
    out = os.path.join(OUT, "libpiko_scene.so")
    tmp = _atomic_out(out)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INC,
          os.path.join(CSRC, "scenes.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_hip():
    out = os.path.join(OUT, "libpiko_hip.so")
    tmp = _atomic_out(out)
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    _run(["/opt/rocm/bin/hipcc"] + HIP_FLAGS + ["-I", INC, "-I", CSRC] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_module():
    out = os.path.join(OUT, "vertex_module.hsaco")
    tmp = out + ".tmp%d" % os.getpid()
    flags = [f for f in HIP_FLAGS if f not in ("-fPIC", "-shared")]
    _run(["/opt/rocm/bin/hipcc", "--genco"] + flags + ["-I", INC, "-I", CSRC,
                                                        os.path.join(CSRC, "module", "vertex_module.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_scene(), build_hip(), build_module()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
