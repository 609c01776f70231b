/*
 * piko/piko.h -- the C++ pipeline surface of the MI355X Piko backend.
 *
 * The user-facing API of Piko (api/include/piko/{pipe,stage,builtinTypes}.h):
 * Stage<binX, binY, threadsPerTile, In, Out> classes with assignBin /
 * schedule / process phases and their directives, PikoPipe, pikoConnect,
 * PikoScreen and PikoArray, ConstantState and MutableState.  A pipe class
 * written against it is compiled by pikoc_hip (--target=HIP), which emits
 * __pikoDefines.h and __pikoCompiledPipe.h: the pipe's allocate / prepare /
 * run_single / run / destroy methods.  When every stage body (read with
 * __PIKOC_DEVICE__ defined) is one the library's hand-written gfx950
 * kernels implement, those methods call the C-ABI of libpiko_hip.so
 * (include/piko_hip.h); otherwise pikoc_hip compiles the bodies themselves
 * for gfx950 (__pikoStages.hip with piko/stage_exec.h, built into
 * lib<Pipe>_stages.so) and the methods call that library.  The pipe's
 * PIKO_STAGE_CODE macro (AUTO, LIBRARY, COMPILED) chooses.
 */
#ifndef PIKO_PIKO_H
#define PIKO_PIKO_H

#include "piko/builtinTypes.h"
#include "piko/pipe.h"
#include "piko/stage.h"

#endif
