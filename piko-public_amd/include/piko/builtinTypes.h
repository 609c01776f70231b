/*
 * piko/builtinTypes.h -- pipeline state and primitive types
 * (api/include/piko/builtinTypes.h:20-55, api/include/piko/cvec.h).
 */
#ifndef PIKO_BUILTIN_TYPES_H
#define PIKO_BUILTIN_TYPES_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "piko_hip.h"

/* Abort with the status text, as the reference's CUDACHECK does
 * (api/include/internal/cudaMacros.h:10-20). */
#define PIKO_CHECK(call)                                                                \
  do {                                                                                  \
    const int piko_rc_ = (call);                                                        \
    if (piko_rc_ != PIKO_OK) {                                                          \
      fprintf(stderr, "%s:%d: %s: %s (%d)\n", __FILE__, __LINE__, #call,               \
              piko_status_string(piko_rc_), piko_rc_);                                  \
      exit(-1);                                                                         \
    }                                                                                   \
  } while (0)

struct cvec2f { float x, y; };
struct cvec3f { float x, y, z; };
struct cvec4f { float x, y, z, w; };
struct cvec2i { int x, y; };

/* ConstantState (builtinTypes.h:20-33): screen size, half size, view-projection
 * (column-major, matmultfloat4) and the vertex lighting constants. */
typedef piko_constant_state ConstantState;

/* MutableState (builtinTypes.h:35-40).  The depth buffer lives on the device
 * and is reset by the pipe each frame; when zBuffer is set, run_single copies
 * the frame's depth (screenSizeX * screenSizeY floats) into it. */
struct MutableState {
  float *zBuffer = nullptr;
};

/* A framebuffer pixel emitted to the PikoScreen (builtinTypes.h:43-56). */
struct Pixel {
  cvec2i pos;
  uint32_t color;
};

struct Primitive {};

#endif
