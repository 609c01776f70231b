"""HIP backend of pikoc_hip: the --target=HIP counterpart of pikoc's
PikoBackend (pikoc/include/Backend/PikoBackend.hpp:14-46), with the same
emit methods.

pikoc_hip takes two stage graphs:
  raster  a streaming stage -> a binned stage that emits Pixels to the
          PikoScreen (the raster sample: VertexShade -> Raster)
  Reyes   a stage looping on itself through out port 1, out port 0 ->
          a streaming stage -> a binned stage emitting samples to the
          PikoScreen (Split -> Dice -> Shade)
and binds a pipe one of two ways (the pipe's PIKO_STAGE_CODE macro: AUTO,
the default, LIBRARY or COMPILED):
  library   the stage bodies are ones the library's hand-written gfx950
            kernels implement, recognised by the fingerprint of their
            normalised source (stage_library.py), with their settings
            (square raster bins of 8-64 px, GORAUD); the generated host code
            calls the C-ABI of libpiko_hip.so (include/piko_hip.h) and
            emitDeviceCode writes the kernel map of the library kernels the
            pipe launches.
  compiled  any other stage bodies (AUTO), or every pipe (COMPILED):
            emitDeviceCode writes __pikoStages.hip, which includes the pipe's
            stage headers with every function made callable from device
            code, plus piko/stage_exec.h (the schedule that runs them in the
            serial order), and builds it with hipcc into
            lib<Pipe>_stages.so, as pikoc compiles stage bodies to PTX
            (pikoc/src/pikoc.cpp:69-318, PTXBackend.cpp:356-647).  The
            generated host code calls that library's C-ABI.
PIKO_STAGE_CODE=LIBRARY refuses a pipe the library does not implement; a
graph of another shape or an invalid setting is refused in every mode
(BackendError; pikoc_hip exits with status 5 and writes nothing).
"""
import os

from . import stage_library
from .summary import PipeSummary

# raster paths the generated pipe may select with PIKO_RASTER_PATH
# (piko_raster_options.raster_mode)
RASTER_PATHS = {"BINNED": 1, "DIRECT": 2, "TILED": 3, "QUEUE": 4}


class BackendError(Exception):
    pass


class NotInLibrary(BackendError):
    """The pipe is not one the kernel library implements (compiled instead
    unless PIKO_STAGE_CODE=LIBRARY)."""


STAGE_CODE = ("AUTO", "LIBRARY", "COMPILED")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# the flags of the library build: gfx950, no contraction into FMAs (the
# serial order's rounding), no packed FP32 instructions
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _macro(ps, name, default, kind=float):
    from .frontend import eval_number
    if name not in ps.macros:
        return default
    return kind(eval_number(ps.macros[name], ps.macros))


def _field(fields, ctype, count, nth=0):
    hits = [f for f in (fields or []) if f[0] in ctype and (f[2] or "").strip() == str(count)]
    return hits[nth][1] if len(hits) > nth else None


class HIPBackend:
    TARGET = "HIP"

    def __init__(self, ps: PipeSummary, path=None):
        self.ps = ps
        self.path_in = path
        self.shape = self._shape()
        mode = ps.macros.get("PIKO_STAGE_CODE", "AUTO").strip()
        if mode not in STAGE_CODE:
            raise BackendError("PIKO_STAGE_CODE must be one of %s" % ", ".join(STAGE_CODE))
        self.library_miss = None
        if mode != "COMPILED":
            try:
                self.kind = self._library()
                return
            except NotInLibrary as e:
                if mode == "LIBRARY":
                    raise
                self.library_miss = str(e)
        self.kind = "compiled"
        self._compiled()

    @staticmethod
    def getTargetTriple():
        return "amdgcn-amd-amdhsa"

    def _shape(self):
        ps = self.ps
        st = ps.order
        if len(st) == 2 and not ps.has_loop and st[0].next_by_port == {0: [st[1]]} and not st[1].next:
            return "raster"
        if len(st) == 3:
            split, dice, shade = st
            if (split.loop and split.next_by_port.get(1) == [split] and split.next_by_port.get(0) == [dice]
                    and dice.next_by_port == {0: [shade]} and not shade.next):
                return "reyes"
        raise BackendError("stage graph of %s is not one pikoc_hip takes (a stage feeding a binned stage that "
                           "writes the PikoScreen, or Split(loop on port 1) -> Dice -> Shade)" % ps.name)

    # -- library matching -------------------------------------------------
    def _library(self):
        ps = self.ps
        st = ps.order
        if self.shape == "raster":
            if not (st[1].binsize[0] == st[1].binsize[1] and st[1].binsize[0] in (8, 16, 32, 64)):
                raise NotInLibrary("%s: the kernel library takes square bins of 8, 16, 32 or 64 pixels"
                                   % st[1].name)
            self.T = st[1].binsize[0]
            self.S = int(_macro(ps, "PIKO_SUBPIXEL_BITS", 4, int))
            if not 1 <= self.S <= 8:
                raise BackendError("PIKO_SUBPIXEL_BITS must be 1..8")
            self.pretransformed = "VTX_PRETRANSFORM" in ps.macros
            path = ps.macros.get("PIKO_RASTER_PATH", "DIRECT").strip()
            if path not in RASTER_PATHS:
                raise BackendError("PIKO_RASTER_PATH must be one of %s" % ", ".join(sorted(RASTER_PATHS)))
            self.path = path
            if "GORAUD" not in ps.macros:
                raise NotInLibrary("%s: shading from interpolated per-pixel normals (GORAUD not defined) is "
                                   "not implemented by the HIP kernel library" % st[1].name)
            self._bodies([("raster.vertex", st[0], "pretransformed" if self.pretransformed else "world"),
                          ("raster.raster", st[1], "gouraud")])
            self.f_pos = _field(ps.input_fields, ("cvec3f",), 3, 0)
            self.f_attr = (_field(ps.input_fields, ("uint32_t", "unsigned", "unsignedint"), 3, 0)
                           if self.pretransformed else _field(ps.input_fields, ("cvec3f",), 3, 1))
            if not self.f_pos or not self.f_attr:
                raise NotInLibrary("input %s needs cvec3f[3] positions and %s" % (
                    ps.input_type, "uint32_t[3] packed colours" if self.pretransformed else "cvec3f[3] normals"))
            return "raster"
        self.threshold = float(_macro(ps, "THRESHOLD_PIXELS", 4.0))
        self.max_splits = int(_macro(ps, "PIKO_REYES_MAX_SPLITS", 20, int))
        self.dice = int(_macro(ps, "PIKO_REYES_DICE", 0, int))
        self.spp = int(_macro(ps, "PIKO_REYES_SPP", 1, int))
        self.disp = float(_macro(ps, "PIKO_REYES_DISPLACEMENT", 0.0))
        # sample compositing: 0 = auto, 1 = global atomics, 2 = tiled
        # (piko_reyes_options.sampling); the tiled path's tile edge
        self.sampling = int(_macro(ps, "PIKO_REYES_SAMPLING", 0, int))
        self.tile = int(_macro(ps, "PIKO_REYES_TILE", 0, int))
        if self.sampling not in (0, 1, 2) or self.tile not in (0, 16, 32):
            raise BackendError("PIKO_REYES_SAMPLING must be 0, 1 or 2 and PIKO_REYES_TILE 0, 16 or 32")
        self._bodies([("reyes.split", st[0], "split"), ("reyes.dice", st[1], "dice"),
                      ("reyes.shade", st[2], "shade")])
        self.f_cp = _field(ps.input_fields, ("cvec4f",), 16, 0)
        if not self.f_cp:
            raise NotInLibrary("input %s needs cvec4f CP[16]" % ps.input_type)
        return "reyes"

    def _bodies(self, checks):
        """Every (role, stage, expected variant) must carry a recognised body."""
        for role, stage, want in checks:
            variant, diff = stage_library.match(role, stage)
            if diff:
                what = ("code outside assignBin / schedule / process" if diff == ["class"]
                        else " and ".join("%s.%s" % (stage.name, ph) for ph in diff))
                raise NotInLibrary("%s of %s is not a body the HIP kernel library implements (compared with its "
                                   "%s stage)" % (what, stage.name, variant))
            if variant != want:
                raise NotInLibrary("%s: the %s body does not match the pipe's settings (%s expected)"
                                   % (stage.name, variant, want))

    # -- compiled stage code ---------------------------------------------------
    def _compiled(self):
        ps = self.ps
        if not self.path_in:
            raise BackendError("compiling stage code needs the pipe header's path")
        self.classes = [st.type for st in ps.order]
        self.prefix = "%s_stages" % ps.name
        self.spp = int(_macro(ps, "PIKO_REYES_SPP", 1, int)) if self.shape == "reyes" else 1
        self.max_levels = int(_macro(ps, "PIKO_MAX_LEVELS", 64, int))
        if not 1 <= self.spp <= 64 or self.max_levels < 1:
            raise BackendError("PIKO_REYES_SPP must be 1..64 and PIKO_MAX_LEVELS at least 1")

    def kernel_map(self):
        """(phase, library kernels, launched every frame) in launch order:
        the kernels the generated pipe runs (kernel names without template
        arguments are what a kernel trace shows)."""
        if self.kind == "compiled":
            return self._compiled_kernel_map()
        if self.kind == "raster":
            a, b = self.ps.order
            if self.path == "DIRECT":
                return [
                    ("kernel0 + %s.process + %s.assignBin + %s.process of the triangles with at most 16,384 "
                     "bounding-box pixels: per-pixel 64-bit visibility atomics" % (a.name, b.name, b.name),
                     ["k_vs_direct"], True),
                    ("%s.process of the larger triangles, one workgroup each (frames that have any)" % b.name,
                     ["k_large_raster"], False),
                    ("depth test winner -> Gouraud colour -> PikoScreen", ["k_resolve"], True),
                ]
            if self.path == "TILED":
                return [
                    ("kernel0 + %s.process + %s.assignBin into per-chunk tile tables" % (a.name, b.name),
                     ["k_tile_bin"], True),
                    ("%s.process per tile in LDS -> PikoScreen (tile = max(T, 16), doubled until the grid "
                     "fits)" % b.name, ["k_tile_raster"], True),
                ]
            if self.path == "QUEUE":
                return [
                    ("kernel0 + %s.process + %s.assignBin into per-tile ranges" % (a.name, b.name),
                     ["k_qbin"], True),
                    ("%s.process: persistent workgroups pull tile parts, LDS depth test -> PikoScreen"
                     % b.name, ["k_qraster"], True),
                    ("merge of tiles split into parts (frames that have any)", ["k_qmerge"], False),
                ]
            return [
                ("kernel0 + %s.process + emit -> %s.assignBin (bin walk, counts)" % (a.name, b.name),
                 ["k_vs_count"], True),
                ("queue offsets", ["k_scan_tiles"], True),
                ("%s.assignBin emits (triangle, bin) entries" % b.name, ["k_emit_pairs"], True),
                ("bin queues in submission order (8-bit digit passes)",
                 ["k_sort_count", "k_sort_rows", "k_sort_scatter"], True),
                ("queue ranges", ["k_bounds"], True),
                ("%s.process (coverage, depth, Gouraud) -> PikoScreen" % b.name, ["k_raster"], True),
            ]
        s, d, sh = self.ps.order
        return [
            ("kernel0 + %s.process looping on port 1 until empty" % s.name, ["k_reyes_split"], True),
            ("emission order of the leaves", ["k_reyes_levels", "k_reyes_rank"], True),
            ("%s.process + emit -> %s.assignBin + %s.process" % (d.name, sh.name, sh.name),
             ["k_reyes_dice"] if self.sampling != 2 else
             ["k_reyes_dice_bin", "k_reyes_tile_count", "k_reyes_tile_scan", "k_reyes_tile_scatter"], True),
            ("PikoScreen resolve", ["k_reyes_resolve"] if self.sampling != 2 else ["k_reyes_tile_sample"], True),
        ]

    # -- emitters ------------------------------------------------------------
    def emitDefines(self, out):
        out.write("// generated by pikoc_hip --target=HIP (%s) from %s\n" % (self.getTargetTriple(),
                                                                           self.ps.filename))
        out.write("#define __PIKOC_HIP__\n\n")
        if self.kind == "compiled":
            out.write("// the pipe runs its own stage bodies, compiled for gfx950 into lib%s.so\n"
                      "#define __PIKOC_HIP_STAGE_CODE__\n\n" % self.prefix)
            out.write("// device-side members: the pipe object of lib%s.so\n" % self.prefix)
            out.write("#define __PIKO_DEVICE_MEMBERS__ void *piko_hip_pipe_ = nullptr;\n")
            return True
        member = "piko_raster_pipe" if self.kind == "raster" else "piko_reyes_pipe"
        out.write("// device-side members: the pipe object of libpiko_hip.so\n")
        out.write("#define __PIKO_DEVICE_MEMBERS__ %s *piko_hip_pipe_ = nullptr;\n" % member)
        return True

    def _head(self, out):
        lib = ("lib%s.so (built from __pikoStages.hip)" % self.prefix if self.kind == "compiled"
               else "libpiko_hip.so (include/piko_hip.h)")
        out.write("// generated by pikoc_hip --target=HIP from %s: host methods of %s over the\n"
                  "// C-ABI of %s\n" % (self.ps.filename, self.ps.name, lib))
        # the stage-code build includes the pipe header for its stages only
        out.write("#ifndef __PIKOC_STAGE_CODE__\n")
        out.write("#ifndef PIKO_%s_COMPILED_H\n#define PIKO_%s_COMPILED_H\n\n" % (self.ps.name, self.ps.name))
        out.write("#include <string.h>\n\n#include <vector>\n\n#include \"piko_hip.h\"\n\n")

    def emitRunFunc(self, out):
        ps = self.ps
        self._head(out)
        if self.kind == "compiled":
            self._compiled_abi(out)
        elif self.kind == "raster":
            out.write("inline int piko_fetch_%s(void *pipe, std::vector<unsigned> &out) {\n"
                      "  return piko_raster_get_data((piko_raster_pipe *)pipe, out.data());\n}\n\n" % ps.name)
        else:
            out.write("// PikoScreen words of the resolved frame: cvec2uintcolor of the FP32 colour\n"
                      "// (split.pikostage:16-18; alpha 1)\n"
                      "inline int piko_fetch_%s(void *pipe, std::vector<unsigned> &out) {\n"
                      "  std::vector<float> rgba(out.size() * 4);\n"
                      "  const int rc = piko_reyes_get_color((piko_reyes_pipe *)pipe, rgba.data());\n"
                      "  if (rc != PIKO_OK) return rc;\n"
                      "  for (size_t i = 0; i < out.size(); ++i) {\n"
                      "    const float *c = &rgba[4 * i];\n"
                      "    out[i] = ((unsigned)(c[3] * 255.0f) << 24) | ((unsigned)(c[2] * 255.0f) << 16) |\n"
                      "             ((unsigned)(c[1] * 255.0f) << 8) | (unsigned)(c[0] * 255.0f);\n"
                      "  }\n  return PIKO_OK;\n}\n\n" % ps.name)
        out.write("// one frame: allocate + prepare + run_single; the pipe stays allocated for\n"
                  "// further prepare / run_single calls until destroy()\n")
        out.write("inline void %s::run(%s &constState, %s &mutableState, %s *input, int count) {\n"
                  "  allocate(constState, mutableState, input, count);\n  prepare();\n  run_single();\n}\n\n"
                  % (ps.name, ps.const_state_type, ps.mutable_state_type, ps.input_type))
        return True

    def emitAllocateFunc(self, out):
        ps = self.ps
        out.write("inline void %s::allocate(%s &constState, %s &mutableState, %s *input, int count) {\n"
                  % (ps.name, ps.const_state_type, ps.mutable_state_type, ps.input_type))
        out.write("  constState_ = &constState;\n  mutableState_ = &mutableState;\n  count_ = count;\n"
                  "  h_input.copyData(input, count);\n")
        if self.kind == "compiled":
            out.write("  if (%s_input_size() != (int)sizeof(%s)) {\n"
                      "    fprintf(stderr, \"lib%s.so was built for another %s layout\\n\");\n"
                      "    exit(-1);\n  }\n" % (self.prefix, ps.input_type, self.prefix, ps.input_type))
            out.write("  PIKO_CHECK(%s_create(&piko_hip_pipe_, &constState, h_input.data(), count));\n"
                      % self.prefix)
        elif self.kind == "raster":
            attr_t = "uint32_t" if self.pretransformed else "float"
            attr_n = 3 if self.pretransformed else 9
            out.write("  std::vector<float> pos((size_t)count * 9);\n")
            out.write("  std::vector<%s> attr((size_t)count * %d);\n" % (attr_t, attr_n))
            out.write("  for (int i = 0; i < count; ++i)\n    for (int v = 0; v < 3; ++v) {\n")
            out.write("      pos[9 * (size_t)i + 3 * v + 0] = input[i].%s[v].x;\n" % self.f_pos)
            out.write("      pos[9 * (size_t)i + 3 * v + 1] = input[i].%s[v].y;\n" % self.f_pos)
            out.write("      pos[9 * (size_t)i + 3 * v + 2] = input[i].%s[v].z;\n" % self.f_pos)
            if self.pretransformed:
                out.write("      attr[3 * (size_t)i + v] = input[i].%s[v];\n" % self.f_attr)
            else:
                for k, c in enumerate("xyz"):
                    out.write("      attr[9 * (size_t)i + 3 * v + %d] = input[i].%s[v].%s;\n" % (k, self.f_attr, c))
            out.write("    }\n")
            out.write("  piko_raster_options opt;\n  memset(&opt, 0, sizeof(opt));\n")
            out.write("  opt.bin_size = %d;       // %s bins\n" % (self.T, ps.order[1].name))
            out.write("  opt.subpixel_bits = %d;\n" % self.S)
            out.write("  opt.input_kind = %s;\n" % ("PIKO_INPUT_PRETRANSFORMED" if self.pretransformed
                                                  else "PIKO_INPUT_WORLD"))
            out.write("  opt.raster_mode = PIKO_RASTER_%s;\n" % self.path)
            out.write("  PIKO_CHECK(piko_raster_allocate(&piko_hip_pipe_, &constState, pos.data(), attr.data(), "
                      "count, &opt));\n")
        else:
            out.write("  std::vector<float> cp((size_t)count * 64);\n")
            out.write("  for (int i = 0; i < count; ++i)\n    for (int k = 0; k < 16; ++k) {\n")
            for k, c in enumerate("xyzw"):
                out.write("      cp[64 * (size_t)i + 4 * k + %d] = input[i].%s[k].%s;\n" % (k, self.f_cp, c))
            out.write("    }\n")
            out.write("  piko_reyes_options opt;\n  memset(&opt, 0, sizeof(opt));\n")
            out.write("  opt.threshold = %rf;\n  opt.max_splits = %d;\n  opt.dice = %d;\n  opt.spp = %d;\n"
                      "  opt.displacement = %rf;\n" % (self.threshold, self.max_splits, self.dice, self.spp,
                                                     self.disp))
            if self.sampling or self.tile:
                out.write("  opt.sampling = %d;\n  opt.tile = %d;\n" % (self.sampling, self.tile))
            out.write("  PIKO_CHECK(piko_reyes_allocate(&piko_hip_pipe_, &constState, cp.data(), count, &opt));\n")
        out.write("  pikoScreen.bind(piko_hip_pipe_, &piko_fetch_%s,\n"
                  "                  (size_t)constState.screenSizeX * (size_t)constState.screenSizeY);\n}\n\n"
                  % ps.name)
        return True

    def emitPrepareFunc(self, out):
        fn = {"raster": "piko_raster_prepare", "reyes": "piko_reyes_prepare"}.get(self.kind,
                                                                             "%s_prepare" % getattr(self, "prefix", ""))
        out.write("inline void %s::prepare() { PIKO_CHECK(%s(piko_hip_pipe_, constState_)); }\n\n"
                  % (self.ps.name, fn))
        return True

    def emitRunSingleFunc(self, out):
        run, depth = (("%s_run" % self.prefix, "%s_get_depth" % self.prefix) if self.kind == "compiled" else
                      ("piko_%s_run_single" % self.kind, "piko_%s_get_depth" % self.kind))
        out.write("inline void %s::run_single() {\n  PIKO_CHECK(%s(piko_hip_pipe_));\n"
                  "  if (mutableState_ && mutableState_->zBuffer)\n"
                  "    PIKO_CHECK(%s(piko_hip_pipe_, mutableState_->zBuffer));\n}\n\n"
                  % (self.ps.name, run, depth))
        return True

    def emitDestroyFunc(self, out):
        ps = self.ps
        fn = "%s_destroy" % self.prefix if self.kind == "compiled" else "piko_%s_destroy" % self.kind
        out.write("inline void %s::destroy() {\n  if (piko_hip_pipe_) {\n"
                  "    PIKO_CHECK(%s(piko_hip_pipe_));\n    piko_hip_pipe_ = nullptr;\n  }\n}\n\n"
                  % (ps.name, fn))
        # one frame with the time of each launch (names as a kernel trace
        # shows them), the same call for either binding
        prof = ("%s_profile(p.piko_hip_pipe_, ms, names, cap, nk, frame_ms)" % self.prefix
                if self.kind == "compiled" else
                "piko_%s_profile_frames(p.piko_hip_pipe_, 1, ms, names, cap, nk, frame_ms)" % self.kind)
        out.write("inline int piko_profile_frame(%s &p, float *ms, const char **names, int cap, int *nk,\n"
                  "                              float *frame_ms) {\n  return %s;\n}\n\n" % (ps.name, prof))
        out.write("#endif\n#endif\n")
        return True

    def emitDeviceCode(self, filename):
        """Writes the kernel map of the pipe to `filename`.kernels and, for
        compiled stage code, the device source __pikoStages.hip beside it."""
        lib = "lib%s.so" % self.prefix if self.kind == "compiled" else "libpiko_hip.so"
        with open(filename + ".kernels", "w") as f:
            f.write("# %s: phases -> %s kernels launched (%s); '?' = only in frames "
                    "that need it\n" % (self.ps.name, lib, self.getTargetTriple()))
            for phase, kernels, always in self.kernel_map():
                f.write("%-70s %s%s\n" % (phase, " ".join(kernels), "" if always else " ?"))
        if self.kind == "compiled":
            with open(os.path.join(os.path.dirname(filename), "__pikoStages.hip"), "w") as f:
                self._stage_source(f)
        return True

    # -- compiled stage code: emitters ------------------------------------------------
    def _exec_type(self):
        if self.shape == "raster":
            return "pikox::RasterExec<%s, %s>" % tuple(self.classes)
        return "pikox::ReyesExec<%s, %s, %s>" % tuple(self.classes)

    def _compiled_abi(self, out):
        """Declarations of the stage library's C-ABI and the PikoScreen fetch."""
        ps, p = self.ps, self.prefix
        out.write('extern "C" {\n')
        out.write("int %s_input_size(void);\n" % p)
        out.write("int %s_create(void **pipe, const piko_constant_state *cs, const void *input, int count);\n" % p)
        out.write("int %s_prepare(void *pipe, const piko_constant_state *cs);\n" % p)
        out.write("int %s_run(void *pipe);\n" % p)
        if self.shape == "raster":
            out.write("int %s_get_color(void *pipe, unsigned *words);\n" % p)
        else:
            out.write("int %s_get_color(void *pipe, float *rgba);\n" % p)
        out.write("int %s_get_depth(void *pipe, float *depth);\n" % p)
        out.write("int %s_profile(void *pipe, float *ms, const char **names, int cap, int *nk, float *frame_ms);\n"
                  % p)
        out.write("int %s_destroy(void *pipe);\n}\n\n" % p)
        if self.shape == "raster":
            out.write("inline int piko_fetch_%s(void *pipe, std::vector<unsigned> &out) {\n"
                      "  return %s_get_color(pipe, out.data());\n}\n\n" % (ps.name, p))
        else:
            out.write("// PikoScreen words of the resolved frame: cvec2uintcolor of the FP32 colour (alpha 1)\n"
                      "inline int piko_fetch_%s(void *pipe, std::vector<unsigned> &out) {\n"
                      "  std::vector<float> rgba(out.size() * 4);\n"
                      "  const int rc = %s_get_color(pipe, rgba.data());\n"
                      "  if (rc != PIKO_OK) return rc;\n"
                      "  for (size_t i = 0; i < out.size(); ++i) {\n"
                      "    const float *c = &rgba[4 * i];\n"
                      "    out[i] = ((unsigned)(c[3] * 255.0f) << 24) | ((unsigned)(c[2] * 255.0f) << 16) |\n"
                      "             ((unsigned)(c[1] * 255.0f) << 8) | (unsigned)(c[0] * 255.0f);\n"
                      "  }\n  return PIKO_OK;\n}\n\n" % (ps.name, p))

    def _stage_source(self, out):
        ps, p = self.ps, self.prefix
        out.write("// __pikoStages.hip -- generated by pikoc_hip --target=HIP (%s) from %s: the\n"
                  "// stage bodies of %s compiled for gfx950 and run by the schedule of\n"
                  "// piko/stage_exec.h, behind the C-ABI the generated host methods call.\n"
                  % (self.getTargetTriple(), ps.filename, ps.name))
        if self.library_miss:
            out.write("// Not the kernel library's: %s.\n" % self.library_miss)
        out.write("#define __PIKOC_DEVICE__ 1\n#define __PIKOC_STAGE_CODE__ 1\n\n")
        # what the stage headers may include stays host-only: only their own
        # declarations become callable from device code
        for h in ("hip/hip_runtime.h", "math.h", "stdint.h", "stdio.h", "stdlib.h", "string.h"):
            out.write("#include <%s>\n" % h)
        out.write("\n")
        for h in ("algorithm", "functional", "limits", "type_traits", "utility", "vector"):
            out.write("#include <%s>\n" % h)
        out.write('\n#include "piko_hip.h"\n\n')
        out.write("#pragma clang force_cuda_host_device begin\n")
        out.write('#include "%s"\n' % ps.filename)
        out.write("#pragma clang force_cuda_host_device end\n\n")
        out.write('#include "piko/stage_exec.h"\n\n')
        out.write("using PipeExec = %s;\n\n" % self._exec_type())
        out.write('extern "C" {\n\n')
        out.write("int %s_input_size(void) { return (int)sizeof(%s); }\n\n" % (p, ps.input_type))
        create_args = ("cs, static_cast<const PipeExec::In *>(input), count" if self.shape == "raster" else
                       "cs, static_cast<const PipeExec::In *>(input), count, %d, %d" % (self.spp, self.max_levels))
        out.write("int %s_create(void **pipe, const piko_constant_state *cs, const void *input, int count) {\n"
                  "  if (!pipe) return PIKO_ERR_INVALID;\n"
                  "  PipeExec *e = new PipeExec();\n"
                  "  const int rc = e->create(%s);\n"
                  "  if (rc != PIKO_OK) {\n    delete e;\n    return rc;\n  }\n"
                  "  *pipe = e;\n  return PIKO_OK;\n}\n\n" % (p, create_args))
        out.write("int %s_prepare(void *pipe, const piko_constant_state *cs) {\n"
                  "  return pipe ? static_cast<PipeExec *>(pipe)->prepare(cs) : PIKO_ERR_INVALID;\n}\n\n" % p)
        out.write("int %s_run(void *pipe) { return pipe ? static_cast<PipeExec *>(pipe)->run() : PIKO_ERR_INVALID; }\n\n"
                  % p)
        ctype = "unsigned *words" if self.shape == "raster" else "float *rgba"
        cname = "words" if self.shape == "raster" else "rgba"
        out.write("int %s_get_color(void *pipe, %s) {\n"
                  "  return pipe && %s ? static_cast<PipeExec *>(pipe)->get_color(%s) : PIKO_ERR_INVALID;\n}\n\n"
                  % (p, ctype, cname, cname))
        out.write("int %s_get_depth(void *pipe, float *depth) {\n"
                  "  return pipe && depth ? static_cast<PipeExec *>(pipe)->get_depth(depth) : PIKO_ERR_INVALID;\n"
                  "}\n\n" % p)
        out.write("int %s_profile(void *pipe, float *ms, const char **names, int cap, int *nk, float *frame_ms) {\n"
                  "  return pipe ? static_cast<PipeExec *>(pipe)->profile(ms, names, cap, nk, frame_ms)"
                  " : PIKO_ERR_INVALID;\n}\n\n" % p)
        out.write("int %s_destroy(void *pipe) {\n  if (!pipe) return PIKO_ERR_INVALID;\n"
                  "  delete static_cast<PipeExec *>(pipe);\n  return PIKO_OK;\n}\n\n}  // extern \"C\"\n" % p)

    def _compiled_kernel_map(self):
        st = self.ps.order
        rows = []
        if self.shape == "raster":
            a, b = st
            rows.append(("screen cleared to the clear colour, depth 1.0", ["k_stage_fill"], True))
            rows.append(("%s.assignBin + %s.process, a thread per primitive; emits appended with (primitive, "
                         "emit number)" % (a.name, a.name), ["k_stage_stream"], True))
        else:
            sp, di, b = st
            rows.append(("screen samples cleared, depth 1.0", ["k_stage_fill", "k_stage_clear_samples"], True))
            rows.append(("%s.assignBin + %s.process level by level (port 1: next level, port 0: leaves); "
                         "then %s.assignBin + %s.process, a thread per primitive" % (sp.name, sp.name, di.name,
                                                                                    di.name),
                         ["k_stage_stream"], True))
        rows.append(("emits put back in primitive order", ["k_stage_scan", "k_stage_order"], True))
        rows.append(("scan levels past 2,048 counts", ["k_stage_scan_add"], False))
        rows.append(("%s.assignBin per primitive: (bin, primitive) in emission order" % b.name,
                     ["k_stage_assign", "k_stage_order_bins"], True))
        rows.append(("stable radix sort by bin, queue bounds", ["k_stage_sort_hist", "k_stage_sort_scatter",
                                                                 "k_stage_bin_bounds"], True))
        rows.append(("%s.process: a thread per bin over its queue in order -> PikoScreen" % b.name,
                     ["k_stage_bins"], True))
        if self.shape == "reyes":
            rows.append(("PikoScreen: box filter over each pixel's samples", ["k_stage_resolve"], True))
        return rows

    def build_stage_code(self, outdir, include_dirs=(), defines=None):
        """hipcc on __pikoStages.hip -> outdir/lib<Pipe>_stages.so."""
        import subprocess
        here = os.path.dirname(os.path.abspath(__file__))
        pkg_inc = os.path.join(os.path.dirname(here), "include")
        abi_inc = os.path.join(os.path.dirname(os.path.dirname(here)), "include")
        src = os.path.join(outdir, "__pikoStages.hip")
        lib = os.path.join(outdir, "lib%s.so" % self.prefix)
        cmd = [HIPCC] + HIPCC_FLAGS + ["-I", outdir, "-I", os.path.dirname(os.path.abspath(self.path_in))]
        for d in include_dirs:
            cmd += ["-I", d]
        cmd += ["-I", pkg_inc, "-I", abi_inc]
        cmd += ["-D%s=%s" % kv for kv in (defines or {}).items()]
        cmd += [src, "-o", lib + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise BackendError("hipcc could not compile the stage code of %s:\n%s" % (self.ps.name, r.stderr[-4000:]))
        os.replace(lib + ".tmp", lib)
        return lib


def compile_pipe(path, outdir, include_dirs=(), defines=None, log=print, build=True):
    """The pikoc driver (pikoc/src/pikoc.cpp:331-463) for --target=HIP:
    summary, kernel plan, __pikoDefines.h, __pikoCompiledPipe.h, the kernel
    map and, for compiled stage code, __pikoStages.hip and (build=True) its
    library lib<Pipe>_stages.so.  Returns the backend."""
    from .frontend import parse_pipe
    ps = parse_pipe(path, include_dirs, defines)
    if log:
        log(ps.display())
    be = HIPBackend(ps, path)
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, "__pikoDefines.h"), "w") as f:
        be.emitDefines(f)
    with open(os.path.join(outdir, "__pikoCompiledPipe.h"), "w") as f:
        for step in (be.emitRunFunc, be.emitAllocateFunc, be.emitPrepareFunc, be.emitRunSingleFunc,
                     be.emitDestroyFunc):
            if not step(f):
                raise BackendError("unable to emit %s" % step.__name__)
    be.emitDeviceCode(os.path.join(outdir, "__pikoCompiledPipe"))
    if be.kind == "compiled":
        if log:
            log("stage code: %s" % ("compiled (PIKO_STAGE_CODE=COMPILED)" if not be.library_miss else
                                    "compiled; not the kernel library's: %s" % be.library_miss))
        if build:
            be.build_stage_code(outdir, include_dirs, defines)
    return be
