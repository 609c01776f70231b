"""pikoc_hip, the --target=HIP code generator: the pipe summary its frontend
reads from the sample pipes, the kernel plan with pikoc's fusion rules
(PikoSummary.cpp:686-755), the emitted host code, the error paths, and that
the samples compile against the C-ABI."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "piko-public_amd")
INC = os.path.join(PKG, "include")
SAMPLES = os.path.join(PKG, "samples")

from pikoc_hip import BackendError, FrontendError, compile_pipe, parse_pipe  # noqa: E402
from pikoc_hip.frontend import eval_number  # noqa: E402
from pikoc_hip.summary import PipeSummary, StageSummary  # noqa: E402


def raster_summary(defines=None):
    return parse_pipe(os.path.join(SAMPLES, "raster", "rasterPipe.h"), [INC], defines)


def test_raster_summary():
    ps = raster_summary()
    assert ps.name == "RasterPipe"
    assert [s.name for s in ps.order] == ["vertexShader", "raster"]
    vs, ra = ps.order
    assert vs.binsize == (32, 32) and vs.threads_per_tile == 512
    assert ra.binsize == (8, 8) and ra.threads_per_tile == 64
    assert (vs.prim_in, vs.prim_out, ra.prim_in, ra.prim_out) == ("raster_wtri", "raster_stri", "raster_stri", "Pixel")
    assert vs.sched_policy == ra.sched_policy == "LoadBalance" and vs.sched_trivial
    assert vs.assign_policy == "Custom"
    assert vs.next_by_port == {0: [ra]} and ra.prev == [vs]
    assert [s.name for s in ps.drain] == ["raster"]
    assert (ps.const_state_type, ps.mutable_state_type, ps.input_type) == ("ConstantState", "MutableState",
                                                                           "raster_wtri")
    assert "VTX_PRETRANSFORM" in ps.macros
    assert [f[1] for f in ps.input_fields] == ["id", "vertPos", "vertColor"]
    # the reference's plan for this pipe: kernel0 = VS.assignBin over the input,
    # kernel1 = VS.process emitting into Raster.assignBin, kernel2 = Raster.process
    plan = ps.kernel_plan()
    assert [k for k, _ in plan] == ["kernel0", "kernel1", "kernel2"]
    assert plan[1][1] == ["vertexShader.process", "emit[0] -> raster.assignBin"]


def test_world_input_define():
    ps = raster_summary({"RASTER_WORLD_INPUT": "1"})
    assert "VTX_PRETRANSFORM" not in ps.macros
    assert [f[1] for f in ps.input_fields] == ["id", "vertPos", "vertNormal"]


def test_reyes_summary_and_loop():
    ps = parse_pipe(os.path.join(SAMPLES, "reyes", "reyesPipe.h"), [INC])
    assert [s.name for s in ps.order] == ["split", "dice", "shade"]
    split, dice, shade = ps.order
    assert split.loop and ps.has_loop and not dice.loop
    assert split.next_by_port == {0: [dice], 1: [split]}
    assert split.binsize == (0, 0) and dice.binsize == shade.binsize == (64, 64)
    assert [s.name for s in ps.drain] == ["shade"]
    assert [k for k, _ in ps.kernel_plan()] == ["kernel0", "kernel1", "kernel2", "kernel3"]


def stage(name, T=8, sched="LoadBalance", assign="InBin", tile=0, wait="None"):
    s = StageSummary(name, name)
    s.binsize, s.sched_policy, s.assign_policy, s.tile_split_size, s.wait_policy = (T, T), sched, assign, tile, wait
    s.sched_trivial = True
    return s


def test_can_fuse_rules():
    f = PipeSummary.can_fuse
    assert f(stage("a"), stage("b"), [])                                   # LoadBalance + InBin, same bins
    assert not f(stage("a"), stage("b", T=16), [])                         # different bins
    assert not f(stage("a"), stage("b", assign="Custom"), [])              # LoadBalance needs InBin
    assert not f(stage("a"), stage("a"), [])                               # same stage type
    assert f(stage("a", sched="Serialize", assign="Custom"), stage("b", sched="Serialize", assign="Custom"), [])
    assert f(stage("a", sched="DirectMap"), stage("b", sched="DirectMap"), [])
    assert not f(stage("a", sched="DirectMap"), stage("b", sched="LoadBalance"), [])
    assert not f(stage("a", tile=2), stage("b", tile=4), [])
    assert f(stage("a", sched="All"), stage("b", sched="All"), [])
    assert not f(stage("a", sched="All"), stage("b", sched="All", wait="EndStage"), [])


def test_can_fuse_end_stage_dependency():
    # PikoSummary.cpp:707-709: a stage waiting for the end of another stage
    # fuses only once that stage's kernel is done
    ps = PipeSummary("P", "p.h")
    for s in (stage("src", T=16), stage("a"), stage("b", wait="EndStage")):
        ps.add_stage(s)
    ps.stage("b").end_stage_name = "src"
    ps.connect("src", "a", 0, 0)
    ps.connect("a", "b", 0, 0)
    ps.process_links()
    src, a, b = ps.stage("src"), ps.stage("a"), ps.stage("b")
    assert b.end_stage is src
    assert not PipeSummary.can_fuse(a, b, [])       # src not done: no fusion
    assert PipeSummary.can_fuse(a, b, [src])        # resolved dependency
    # src (other bins) has its own kernel before a's, so a and b fuse; a
    # dependency on a stage planned later keeps them apart
    assert [p for _, p in ps.kernel_plan()][2] == ["a.process", "emit[0] -> b.assignBin", "b.process"]
    ps.stage("b").end_stage_name = "b2"
    late = stage("b2")
    ps.add_stage(late)
    ps.connect("b", "b2", 0, 0)
    ps.process_links()
    names = [p for _, p in ps.kernel_plan()]
    assert ["a.process", "emit[0] -> b.assignBin"] in names and any(p[0] == "b.process" for p in names)


def test_fused_plan(tmp_path):
    # two InBin stages under LOAD_BALANCE with equal bins fuse into one kernel
    (tmp_path / "p.h").write_text("""
#include "piko/pipe.h"
struct In : public Primitive { cvec3f vertPos[3]; };
class A : public Stage<16, 16, 64, In, In> {
  void assignBin(In p) { this->assignToBin(p, PREVIOUS_BINS); }
  void schedule(int b) { specifySchedule(LOAD_BALANCE); }
  void process(In p) {}
};
class B : public Stage<16, 16, 64, In, Pixel> {
  void assignBin(In p) { this->assignToBin(p, PREVIOUS_BINS); }
  void schedule(int b) { specifySchedule(LOAD_BALANCE); specifyWait(END_BIN); }
  void process(In p) {}
};
class P : public PikoPipe {
  A a; B b;
 public:
  P() { pikoConnect(a, b, 0, 0); }
  void run(ConstantState &c, MutableState &m, In *in, int n);
  void allocate(ConstantState &c, MutableState &m, In *in, int n);
  void prepare(); void run_single(); void destroy();
};
""")
    ps = parse_pipe(str(tmp_path / "p.h"), [INC])
    a, b = ps.order
    assert a.assign_policy == "InBin" and a.assign_trivial and b.wait_policy == "EndBin" and b.sched_trivial
    plan = ps.kernel_plan()
    assert len(plan) == 2 and plan[1][1] == ["a.process", "emit[0] -> b.assignBin", "b.process"]
    # not bodies the kernel library implements: refused when the library is
    # required, compiled otherwise
    with pytest.raises(BackendError):
        compile_pipe(str(tmp_path / "p.h"), str(tmp_path / "out"), [INC], defines={"PIKO_STAGE_CODE": "LIBRARY"},
                     log=None)
    be = compile_pipe(str(tmp_path / "p.h"), str(tmp_path / "out2"), [INC], log=None, build=False)
    assert be.kind == "compiled" and be.shape == "raster" and be.classes == ["A", "B"]
    assert "pikox::RasterExec<A, B>" in (tmp_path / "out2" / "__pikoStages.hip").read_text()


def test_source_errors(tmp_path):
    (tmp_path / "bad.h").write_text("class P : public PikoPipe { P() {} };\n")
    with pytest.raises(FrontendError):
        parse_pipe(str(tmp_path / "bad.h"), [INC])
    (tmp_path / "bad2.h").write_text("#ifdef X\nclass P : public PikoPipe {};\n")
    with pytest.raises(FrontendError):
        parse_pipe(str(tmp_path / "bad2.h"), [INC])


def test_constant_expressions():
    m = {"LG": "3", "T": "(1 << LG)", "F": "16.0f"}
    assert eval_number("T * 2 + 1", m) == 17
    assert eval_number("F / 4", m) == 4.0
    with pytest.raises(FrontendError):
        eval_number("__import__('os')", m)


def test_emitted_raster_host_code(tmp_path):
    be = compile_pipe(os.path.join(SAMPLES, "raster", "rasterPipe.h"), str(tmp_path), [INC], log=None)
    assert be.kind == "raster" and be.T == 8 and be.S == 8 and be.pretransformed
    h = (tmp_path / "__pikoCompiledPipe.h").read_text()
    for needle in ("inline void RasterPipe::allocate(ConstantState &constState", "opt.bin_size = 8;",
                   "opt.subpixel_bits = 8;", "PIKO_INPUT_PRETRANSFORMED", "piko_raster_prepare(piko_hip_pipe_",
                   "piko_raster_run_single(piko_hip_pipe_)", "piko_raster_destroy(piko_hip_pipe_)",
                   "input[i].vertColor[v]"):
        assert needle in h, needle
    d = (tmp_path / "__pikoDefines.h").read_text()
    assert "#define __PIKOC_HIP__" in d and "piko_raster_pipe *piko_hip_pipe_" in d
    # the path the generated pipe selects, and the kernels that path launches
    assert "opt.raster_mode = PIKO_RASTER_DIRECT;" in h
    k = (tmp_path / "__pikoCompiledPipe.kernels").read_text()
    assert [line.split()[-1] for line in k.splitlines()[1:]] == ["k_vs_direct", "?", "k_resolve"]
    assert "k_large_raster ?" in k and "k_vs_count" not in k


@pytest.mark.parametrize("path,kernels", [("QUEUE", ["k_qbin", "k_qraster", "k_qmerge"]),
                                          ("TILED", ["k_tile_bin", "k_tile_raster"]),
                                          ("BINNED", ["k_vs_count", "k_scan_tiles", "k_emit_pairs", "k_sort_count",
                                                      "k_sort_rows", "k_sort_scatter", "k_bounds", "k_raster"])])
def test_raster_path_setting(tmp_path, path, kernels):
    be = compile_pipe(os.path.join(SAMPLES, "raster", "rasterPipe.h"), str(tmp_path), [INC],
                      defines={"PIKO_RASTER_PATH": path}, log=None)
    assert [k for _, ks, _ in be.kernel_map() for k in ks] == kernels
    assert ("opt.raster_mode = PIKO_RASTER_%s;" % path) in (tmp_path / "__pikoCompiledPipe.h").read_text()
    with pytest.raises(BackendError):
        compile_pipe(os.path.join(SAMPLES, "raster", "rasterPipe.h"), str(tmp_path / "x"), [INC],
                     defines={"PIKO_RASTER_PATH": "FAST"}, log=None)


def copy_sample(tmp_path, sub):
    import shutil
    dst = tmp_path / sub
    shutil.copytree(os.path.join(SAMPLES, sub), dst, ignore=shutil.ignore_patterns("gen_*", "*_sample*"))
    return dst


def run_cli(pipe, outdir, *extra):
    env = dict(os.environ, PYTHONPATH=PKG)
    return subprocess.run([sys.executable, "-m", "pikoc_hip", "--target=HIP", "-I", INC, "-o", str(outdir)]
                          + list(extra) + [str(pipe)], env=env, capture_output=True, text=True)


EDITS = [
    # a changed depth rule in the raster stage's process
    ("raster", "raster.h", "< piko::floatBits(z)) continue;", "<= piko::floatBits(z)) continue;",
     "raster.process"),
    # a changed bin assignment in the vertex stage
    ("raster", "vertexShader.h", "(p.id / 512) % getNumBins()", "(p.id / 256) % getNumBins()",
     "vertexShader.assignBin"),
    # an off-by-one in the dice stage's grid
    ("reyes", "stages.h", "const int lastU = PIKO_REYES_DICE > 0 ? nu - 1 : nu",
     "const int lastU = PIKO_REYES_DICE > 0 ? nu : nu", "dice.process"),
    # a changed helper outside the three phases
    ("reyes", "stages.h", "out[0] = piko::saturate(0.2f + d * 0.1773f);",
     "out[0] = piko::saturate(0.25f + d * 0.1773f);", "code outside"),
]


def edited_pipe(tmp_path, sub, header, old, new):
    src = copy_sample(tmp_path, sub)
    text = (src / header).read_text()
    assert old in text
    (src / header).write_text(text.replace(old, new, 1))
    return src / ("rasterPipe.h" if sub == "raster" else "reyesPipe.h")


@pytest.mark.parametrize("sub,header,old,new,names", EDITS)
def test_edited_stage_body_refused_by_library_mode(tmp_path, sub, header, old, new, names):
    pipe = edited_pipe(tmp_path, sub, header, old, new)
    r = run_cli(pipe, tmp_path / "out", "-D", "PIKO_STAGE_CODE=LIBRARY")
    assert r.returncode == 5, r.stderr
    assert names in r.stderr and "is not a body the HIP kernel library implements" in r.stderr
    assert not (tmp_path / "out").exists(), "nothing is written for a refused pipe"


@pytest.mark.parametrize("sub,header,old,new,names", EDITS)
def test_edited_stage_body_compiled(tmp_path, sub, header, old, new, names):
    # the default binding compiles bodies the library does not implement
    pipe = edited_pipe(tmp_path, sub, header, old, new)
    r = run_cli(pipe, tmp_path / "out", "--no-build")
    assert r.returncode == 0, r.stderr
    assert "stage code: compiled; not the kernel library's" in r.stdout and names in r.stdout
    hip = (tmp_path / "out" / "__pikoStages.hip").read_text()
    assert ('#include "%s"' % pipe.name) in hip and "#pragma clang force_cuda_host_device begin" in hip
    assert ("pikox::RasterExec<VertexShaderStage, RasterStage>" if sub == "raster" else
            "pikox::ReyesExec<SplitStage, DiceStage, ShadeStage>") in hip
    glue = (tmp_path / "out" / "__pikoCompiledPipe.h").read_text()
    assert "_stages_create(&piko_hip_pipe_" in glue and "piko_raster_allocate" not in glue
    kmap = (tmp_path / "out" / "__pikoCompiledPipe.kernels").read_text()
    assert "k_stage_bins" in kmap and "k_qraster" not in kmap


def test_compiled_stage_code_builds_and_exports_its_abi(tmp_path):
    # hipcc builds the generated source for gfx950; the library exports the
    # C-ABI the generated host methods declare (nothing is called: no GPU)
    import ctypes
    pipe = edited_pipe(tmp_path, *EDITS[0][:4])
    r = run_cli(pipe, tmp_path / "out")
    assert r.returncode == 0, r.stderr
    lib = tmp_path / "out" / "libRasterPipe_stages.so"
    assert lib.exists()
    so = ctypes.CDLL(str(lib))
    for fn in ("input_size", "create", "prepare", "run", "get_color", "get_depth", "profile", "destroy"):
        assert hasattr(so, "RasterPipe_stages_" + fn), fn
    # raster_wtri of the pretransformed build: id, 3 positions, 3 colour words
    assert so.RasterPipe_stages_input_size() == 4 + 36 + 12


def test_emptied_process_body(tmp_path):
    src = copy_sample(tmp_path, "raster")
    text = (src / "raster.h").read_text()
    start = text.index("  void process(raster_stri p) {")
    end = text.index("\n  }\n", start) + 4
    (src / "raster.h").write_text(text[:start] + "  void process(raster_stri p) {}\n" + text[end:])
    r = run_cli(src / "rasterPipe.h", tmp_path / "out", "-D", "PIKO_STAGE_CODE=LIBRARY")
    assert r.returncode == 5 and "raster.process" in r.stderr
    r = run_cli(src / "rasterPipe.h", tmp_path / "out2", "--no-build")
    assert r.returncode == 0 and (tmp_path / "out2" / "__pikoStages.hip").exists()


def test_reformatted_body_accepted(tmp_path):
    # layout and comments are not part of a body: re-indented, re-wrapped and
    # commented stages are still the recognised ones
    src = copy_sample(tmp_path, "raster")
    text = (src / "raster.h").read_text()
    text = text.replace("    ", "\t").replace(" = ", "   =   ").replace("; ", ";\n /* note */ ")
    (src / "raster.h").write_text(text)
    assert run_cli(src / "rasterPipe.h", tmp_path / "out", "-D", "PIKO_STAGE_CODE=LIBRARY").returncode == 0
    assert not (tmp_path / "out" / "__pikoStages.hip").exists()


def test_missing_goraud(tmp_path):
    src = copy_sample(tmp_path, "raster")
    text = (src / "rasterPipe.h").read_text()
    assert "#define GORAUD\n" in text
    (src / "rasterPipe.h").write_text(text.replace("#define GORAUD\n", ""))
    with pytest.raises(BackendError, match="GORAUD"):
        compile_pipe(str(src / "rasterPipe.h"), str(tmp_path / "out"), [INC],
                     defines={"PIKO_STAGE_CODE": "LIBRARY"}, log=None)
    assert run_cli(src / "rasterPipe.h", tmp_path / "out", "-D", "PIKO_STAGE_CODE=LIBRARY").returncode == 5
    be = compile_pipe(str(src / "rasterPipe.h"), str(tmp_path / "out2"), [INC], log=None, build=False)
    assert be.kind == "compiled" and "GORAUD" in be.library_miss


def test_lit_sample_is_compiled_stage_code(tmp_path):
    be = compile_pipe(os.path.join(SAMPLES, "raster_lit", "rasterPipe.h"), str(tmp_path), [INC], log=None,
                      build=False)
    assert be.kind == "compiled" and be.shape == "raster" and "GORAUD" in be.library_miss


def test_stage_code_setting_and_graph_refusals(tmp_path):
    r = run_cli(os.path.join(SAMPLES, "raster", "rasterPipe.h"), tmp_path / "out", "-D", "PIKO_STAGE_CODE=FAST")
    assert r.returncode == 5 and "PIKO_STAGE_CODE" in r.stderr
    # a three-stage chain without the loop is neither graph pikoc_hip takes
    (tmp_path / "p.h").write_text("""
#include "piko/pipe.h"
struct In : public Primitive { cvec3f vertPos[3]; };
class A : public Stage<16, 16, 64, In, In> {
  void assignBin(In p) { this->assignToBin(p, 0); }
  void schedule(int b) { specifySchedule(LOAD_BALANCE); }
  void process(In p) { this->emit(p, 0); }
};
class B : public Stage<16, 16, 64, In, Pixel> {
  void assignBin(In p) { this->assignToBin(p, 0); }
  void schedule(int b) { specifySchedule(LOAD_BALANCE); }
  void process(In p) {}
};
class P : public PikoPipe {
  A a; A a2; B b;
 public:
  P() { pikoConnect(a, a2, 0, 0); pikoConnect(a2, b, 0, 0); }
  void run(ConstantState &c, MutableState &m, In *in, int n);
  void allocate(ConstantState &c, MutableState &m, In *in, int n);
  void prepare(); void run_single(); void destroy();
};
""")
    for mode in ("AUTO", "COMPILED", "LIBRARY"):
        r = run_cli(tmp_path / "p.h", tmp_path / ("out_" + mode), "-D", "PIKO_STAGE_CODE=" + mode, "--no-build")
        assert r.returncode == 5 and "stage graph" in r.stderr, (mode, r.stderr)


def test_sample_stages_hash_to_library_table():
    from pikoc_hip import stage_library
    cases = [("raster", "rasterPipe.h", {}, ["raster.vertex", "raster.raster"], ["pretransformed", "gouraud"]),
             ("raster", "rasterPipe.h", {"RASTER_WORLD_INPUT": "1"}, ["raster.vertex", "raster.raster"],
              ["world", "gouraud"]),
             ("reyes", "reyesPipe.h", {}, ["reyes.split", "reyes.dice", "reyes.shade"], ["split", "dice", "shade"])]
    for sub, header, defines, roles, variants in cases:
        ps = parse_pipe(os.path.join(SAMPLES, sub, header), [INC], defines)
        for st, role, want in zip(ps.order, roles, variants):
            assert stage_library.match(role, st) == (want, []), (st.name, stage_library.fingerprints(st))


def test_emitted_reyes_host_code(tmp_path):
    be = compile_pipe(os.path.join(SAMPLES, "reyes", "reyesPipe.h"), str(tmp_path), [INC], log=None)
    assert be.kind == "reyes" and be.threshold == 16.0 and be.dice == 8 and be.spp == 4
    h = (tmp_path / "__pikoCompiledPipe.h").read_text()
    assert "piko_reyes_allocate(&piko_hip_pipe_" in h and "opt.dice = 8;" in h and "input[i].CP[k].w" in h
    assert "opt.sampling" not in h  # unset: the zero-filled options pick auto


def test_emitted_reyes_sampling_settings(tmp_path):
    be = compile_pipe(os.path.join(SAMPLES, "reyes", "reyesPipe.h"), str(tmp_path), [INC],
                      defines={"PIKO_REYES_SAMPLING": "2", "PIKO_REYES_TILE": "32"}, log=None)
    assert be.sampling == 2 and be.tile == 32
    h = (tmp_path / "__pikoCompiledPipe.h").read_text()
    assert "opt.sampling = 2;" in h and "opt.tile = 32;" in h
    with pytest.raises(BackendError):
        compile_pipe(os.path.join(SAMPLES, "reyes", "reyesPipe.h"), str(tmp_path), [INC],
                     defines={"PIKO_REYES_TILE": "8"}, log=None)


def test_cli_exit_codes(tmp_path):
    env = dict(os.environ, PYTHONPATH=PKG)
    r = subprocess.run([sys.executable, "-m", "pikoc_hip", "--target=HIP", "-I", INC, "-o", str(tmp_path),
                        os.path.join(SAMPLES, "raster", "rasterPipe.h")], env=env, capture_output=True, text=True)
    assert r.returncode == 0 and "Kernel plan" in r.stdout
    r = subprocess.run([sys.executable, "-m", "pikoc_hip", "--target=PTX", "x.h"], env=env, capture_output=True)
    assert r.returncode == 2


@pytest.mark.parametrize("binary,args", [("raster/raster_sample", []),
                                         ("raster/raster_sample_world", ["--scene", "heightfield"]),
                                         ("reyes/reyes_sample", []),
                                         ("raster/raster_sample_compiled", []),
                                         ("reyes/reyes_sample_compiled", []),
                                         ("raster_lit/raster_lit_sample", ["--scene", "heightfield"]),
                                         ("edge/edge_pipe", [])])
def test_samples_built_and_fail_loudly_without_gpu(binary, args):
    path = os.path.join(SAMPLES, binary) if not binary.startswith("edge/") else \
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "stage_pipes", binary)
    if not os.path.exists(path):
        pytest.skip("samples not built (run __graft_entry__.build())")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the GPU tests run the samples")
    r = subprocess.run([path, "--runs", "1"] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "no HIP device" in r.stderr
