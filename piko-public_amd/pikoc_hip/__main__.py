"""pikoc_hip command line, after pikoc's (pikoc/src/PikocOptions.cpp:35-103):

  python -m pikoc_hip --target=HIP [-I dir]... [-D NAME[=VALUE]]... [-o outdir] [--no-build] pipe.h

writes outdir/__pikoDefines.h, outdir/__pikoCompiledPipe.h and
outdir/__pikoCompiledPipe.kernels and prints the pipe summary and kernel
plan.  When the pipe's stage code is compiled (bodies the kernel library
does not implement, or PIKO_STAGE_CODE=COMPILED) it also writes
outdir/__pikoStages.hip and builds outdir/lib<Pipe>_stages.so with hipcc
(--no-build: the source only).  Exit status: 0 done, 2 usage, 4 output not
writable, 5 a stage graph or setting pikoc_hip does not take, a pipe the
kernel library does not implement under PIKO_STAGE_CODE=LIBRARY, or stage
code hipcc rejects (nothing is written but the sources), 6 source not
understood.

  python -m pikoc_hip --fingerprints [-I dir]... [-D NAME[=VALUE]]... pipe.h

prints the fingerprints of each stage's normalised source, the form of the
recognised-body table in pikoc_hip/stage_library.py.
"""
import argparse
import sys

from .frontend import FrontendError
from .hip_backend import BackendError, compile_pipe


def main(argv=None):
    ap = argparse.ArgumentParser(prog="pikoc_hip")
    ap.add_argument("--target", default="HIP")
    ap.add_argument("-I", dest="include", action="append", default=[])
    ap.add_argument("-D", dest="define", action="append", default=[])
    ap.add_argument("-o", dest="outdir", default=".")
    ap.add_argument("--fingerprints", action="store_true")
    ap.add_argument("--no-build", action="store_true")
    ap.add_argument("pipe")
    a = ap.parse_args(argv)
    if a.target.upper() != "HIP":
        print("pikoc_hip: only --target=HIP is implemented", file=sys.stderr)
        return 2
    defines = {}
    for d in a.define:
        k, _, v = d.partition("=")
        defines[k] = v or "1"
    if a.fingerprints:
        from .frontend import parse_pipe
        from .stage_library import fingerprints
        try:
            ps = parse_pipe(a.pipe, a.include, defines)
        except FrontendError as e:
            print("pikoc_hip: %s" % e, file=sys.stderr)
            return 6
        for st in ps.order:
            print("%s (%s): %s" % (st.name, st.type, fingerprints(st)))
        return 0
    try:
        compile_pipe(a.pipe, a.outdir, a.include, defines, build=not a.no_build)
    except FrontendError as e:
        print("pikoc_hip: %s" % e, file=sys.stderr)
        return 6
    except BackendError as e:
        print("pikoc_hip: %s" % e, file=sys.stderr)
        return 5
    except OSError as e:
        print("pikoc_hip: %s" % e, file=sys.stderr)
        return 4
    return 0


if __name__ == "__main__":
    sys.exit(main())
