"""pikoc_hip command line, after pikoc's (pikoc/src/PikocOptions.cpp:35-103):

  python -m pikoc_hip --target=HIP [-I dir]... [-D NAME[=VALUE]]... [-o outdir] pipe.h

writes outdir/__pikoDefines.h, outdir/__pikoCompiledPipe.h and
outdir/__pikoCompiledPipe.kernels and prints the pipe summary and kernel
plan.  Exit status: 0 done, 2 usage, 4 output not writable, 5 stage graph not
supported by the HIP kernel library, 6 source not understood.
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
    ap.add_argument("pipe")
    a = ap.parse_args(argv)
    if a.target.upper() != "HIP":
        print("pikoc_hip: only --target=HIP is implemented", file=sys.stderr)
        return 2
    defines = {}
    for d in a.define:
        k, _, v = d.partition("=")
        defines[k] = v or "1"
    try:
        compile_pipe(a.pipe, a.outdir, a.include, defines)
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
