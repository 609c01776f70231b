"""HIP backend of pikoc_hip: the --target=HIP counterpart of pikoc's
PikoBackend (pikoc/include/Backend/PikoBackend.hpp:14-46), with the same
emit methods.

The generated host code calls the C-ABI of libpiko_hip.so
(include/piko_hip.h); the device code is the library's hand-written gfx950
kernels.  A pipe is bound to them only when both its stage graph and every
stage body are ones the kernels implement:
  raster  VertexShade stage -> Raster stage with square bins of 8/16/32/64 px
          -> PikoScreen, with GORAUD (vertex colours interpolated)
  Reyes   Split stage looping on itself through out port 1, out port 0 ->
          Dice -> Shade -> PikoScreen
Stage bodies are recognised by the fingerprint of their normalised source
(stage_library.py); any other graph, an edited or unimplemented body, or a
shading variant the kernels lack raises BackendError (pikoc_hip exits with
status 5) and nothing is emitted.  emitDeviceCode writes the kernel map: the
library kernels the generated pipe launches, in launch order.
"""
import os

from . import stage_library
from .summary import PipeSummary

# raster paths the generated pipe may select with PIKO_RASTER_PATH
# (piko_raster_options.raster_mode)
RASTER_PATHS = {"BINNED": 1, "DIRECT": 2, "TILED": 3, "QUEUE": 4}


class BackendError(Exception):
    pass


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

    def __init__(self, ps: PipeSummary):
        self.ps = ps
        self.kind = self._recognise()

    @staticmethod
    def getTargetTriple():
        return "amdgcn-amd-amdhsa"

    # -- library matching -------------------------------------------------
    def _recognise(self):
        ps = self.ps
        st = ps.order
        if (len(st) == 2 and not ps.has_loop and st[0].next_by_port == {0: [st[1]]} and not st[1].next
                and st[1].binsize[0] == st[1].binsize[1] and st[1].binsize[0] in (8, 16, 32, 64)):
            self.T = st[1].binsize[0]
            self.S = int(_macro(ps, "PIKO_SUBPIXEL_BITS", 4, int))
            if not 1 <= self.S <= 8:
                raise BackendError("PIKO_SUBPIXEL_BITS must be 1..8")
            self.pretransformed = "VTX_PRETRANSFORM" in ps.macros
            if "GORAUD" not in ps.macros:
                raise BackendError("%s: shading from interpolated per-pixel normals (GORAUD not defined) is "
                                   "not implemented by the HIP kernel library" % st[1].name)
            self._bodies([("raster.vertex", st[0], "pretransformed" if self.pretransformed else "world"),
                          ("raster.raster", st[1], "gouraud")])
            path = ps.macros.get("PIKO_RASTER_PATH", "DIRECT").strip()
            if path not in RASTER_PATHS:
                raise BackendError("PIKO_RASTER_PATH must be one of %s" % ", ".join(sorted(RASTER_PATHS)))
            self.path = path
            self.f_pos = _field(ps.input_fields, ("cvec3f",), 3, 0)
            self.f_attr = (_field(ps.input_fields, ("uint32_t", "unsigned", "unsignedint"), 3, 0)
                           if self.pretransformed else _field(ps.input_fields, ("cvec3f",), 3, 1))
            if not self.f_pos or not self.f_attr:
                raise BackendError("input %s needs cvec3f[3] positions and %s" % (
                    ps.input_type, "uint32_t[3] packed colours" if self.pretransformed else "cvec3f[3] normals"))
            return "raster"
        if len(st) == 3:
            split, dice, shade = st
            if (split.loop and split.next_by_port.get(1) == [split] and split.next_by_port.get(0) == [dice]
                    and dice.next_by_port == {0: [shade]} and not shade.next):
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
                self.f_cp = _field(ps.input_fields, ("cvec4f",), 16, 0)
                if not self.f_cp:
                    raise BackendError("input %s needs cvec4f CP[16]" % ps.input_type)
                self._bodies([("reyes.split", split, "split"), ("reyes.dice", dice, "dice"),
                              ("reyes.shade", shade, "shade")])
                return "reyes"
        raise BackendError("stage graph of %s is not one the HIP kernel library implements "
                           "(VertexShade -> Raster, or Split(loop) -> Dice -> Shade)" % ps.name)

    def _bodies(self, checks):
        """Every (role, stage, expected variant) must carry a recognised body."""
        for role, stage, want in checks:
            variant, diff = stage_library.match(role, stage)
            if diff:
                what = ("code outside assignBin / schedule / process" if diff == ["class"]
                        else " and ".join("%s.%s" % (stage.name, ph) for ph in diff))
                raise BackendError("%s of %s is not a body the HIP kernel library implements (compared with "
                                   "its %s stage); pikoc_hip does not compile stage code" % (what, stage.name,
                                                                                             variant))
            if variant != want:
                raise BackendError("%s: the %s body does not match the pipe's settings (%s expected)"
                                   % (stage.name, variant, want))

    def kernel_map(self):
        """(phase, library kernels, launched every frame) in launch order:
        the kernels the generated pipe runs (kernel names without template
        arguments are what a kernel trace shows)."""
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
        member = "piko_raster_pipe" if self.kind == "raster" else "piko_reyes_pipe"
        out.write("// generated by pikoc_hip --target=HIP (%s) from %s\n" % (self.getTargetTriple(),
                                                                           self.ps.filename))
        out.write("#define __PIKOC_HIP__\n\n")
        out.write("// device-side members: the pipe object of libpiko_hip.so\n")
        out.write("#define __PIKO_DEVICE_MEMBERS__ %s *piko_hip_pipe_ = nullptr;\n" % member)
        return True

    def _head(self, out):
        out.write("// generated by pikoc_hip --target=HIP from %s: host methods of %s over the\n"
                  "// C-ABI of libpiko_hip.so (include/piko_hip.h)\n" % (self.ps.filename, self.ps.name))
        out.write("#ifndef PIKO_%s_COMPILED_H\n#define PIKO_%s_COMPILED_H\n\n" % (self.ps.name, self.ps.name))
        out.write("#include <string.h>\n\n#include <vector>\n\n#include \"piko_hip.h\"\n\n")

    def emitRunFunc(self, out):
        ps = self.ps
        self._head(out)
        if self.kind == "raster":
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
        if self.kind == "raster":
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
        fn = "piko_raster_prepare" if self.kind == "raster" else "piko_reyes_prepare"
        out.write("inline void %s::prepare() { PIKO_CHECK(%s(piko_hip_pipe_, constState_)); }\n\n"
                  % (self.ps.name, fn))
        return True

    def emitRunSingleFunc(self, out):
        k = self.kind
        out.write("inline void %s::run_single() {\n  PIKO_CHECK(piko_%s_run_single(piko_hip_pipe_));\n"
                  "  if (mutableState_ && mutableState_->zBuffer)\n"
                  "    PIKO_CHECK(piko_%s_get_depth(piko_hip_pipe_, mutableState_->zBuffer));\n}\n\n"
                  % (self.ps.name, k, k))
        return True

    def emitDestroyFunc(self, out):
        out.write("inline void %s::destroy() {\n  if (piko_hip_pipe_) {\n"
                  "    PIKO_CHECK(piko_%s_destroy(piko_hip_pipe_));\n    piko_hip_pipe_ = nullptr;\n  }\n}\n\n"
                  "#endif\n" % (self.ps.name, self.kind))
        return True

    def emitDeviceCode(self, filename):
        """Writes the kernel map of the pipe to `filename`.kernels."""
        with open(filename + ".kernels", "w") as f:
            f.write("# %s: phases -> libpiko_hip.so kernels launched (%s); '?' = only in frames "
                    "that need it\n" % (self.ps.name, self.getTargetTriple()))
            for phase, kernels, always in self.kernel_map():
                f.write("%-70s %s%s\n" % (phase, " ".join(kernels), "" if always else " ?"))
        return True


def compile_pipe(path, outdir, include_dirs=(), defines=None, log=print):
    """The pikoc driver (pikoc/src/pikoc.cpp:331-463) for --target=HIP:
    summary, kernel plan, __pikoDefines.h, __pikoCompiledPipe.h and the
    kernel map.  Returns the backend."""
    from .frontend import parse_pipe
    ps = parse_pipe(path, include_dirs, defines)
    if log:
        log(ps.display())
    be = HIPBackend(ps)
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, "__pikoDefines.h"), "w") as f:
        be.emitDefines(f)
    with open(os.path.join(outdir, "__pikoCompiledPipe.h"), "w") as f:
        for step in (be.emitRunFunc, be.emitAllocateFunc, be.emitPrepareFunc, be.emitRunSingleFunc,
                     be.emitDestroyFunc):
            if not step(f):
                raise BackendError("unable to emit %s" % step.__name__)
    be.emitDeviceCode(os.path.join(outdir, "__pikoCompiledPipe"))
    return be
