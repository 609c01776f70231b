/*
 * piko/device.h -- what a stage body may call (the device-side surface of
 * api/include/piko/{deviceFunctions,math,atomics}.h that the sample stages
 * use).  Stage bodies sit in #ifdef __PIKOC_DEVICE__ blocks, as in the
 * reference's .pikostage files: pikoc_hip reads them with __PIKOC_DEVICE__
 * defined, maps the recognised ones onto the kernel library and compiles
 * the others for gfx950 (the helpers below then run on the device as
 * written).  The pure arithmetic helpers are defined here; the atomic, like
 * the Stage members a body calls, is declared only: piko/stage_exec.h
 * defines it on the device, the tests' host harness on the host.
 */
#ifndef PIKO_DEVICE_H
#define PIKO_DEVICE_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "piko/builtinTypes.h"

namespace piko {

inline int imin(int a, int b) { return a < b ? a : b; }
inline int imax(int a, int b) { return a > b ? a : b; }
inline float fmin2(float a, float b) { return b < a ? b : a; }
inline float fmax2(float a, float b) { return a < b ? b : a; }
inline float saturate(float v) { return v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v); }

inline int32_t floatBits(float f) {
  int32_t i;
  memcpy(&i, &f, 4);
  return i;
}

/* column-major 4x4 matrix times (x, y, z, w) */
inline cvec4f transform4(const float m[16], cvec4f v) {
  cvec4f o;
  o.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  o.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  o.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  o.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return o;
}

/* (int)(v * 2^S), clamped to +-2^29 first so every input converts */
inline int32_t toFixed(float v, int S) {
  float s = v * (float)(1 << S);
  s = s > 536870912.0f ? 536870912.0f : s;
  s = s < -536870912.0f ? -536870912.0f : s;
  return (int32_t)s;
}

/* RGB in [0,1] to the PikoScreen word: truncating, alpha 255 */
inline uint32_t packABGR(float r, float g, float b) {
  return (255u << 24) | ((unsigned)(b * 255.0f) << 16) | ((unsigned)(g * 255.0f) << 8) | (unsigned)(r * 255.0f);
}

inline cvec3f unpackABGR(uint32_t c) {
  cvec3f o;
  o.x = (float)(c & 255u) / 255.0f;
  o.y = (float)((c >> 8) & 255u) / 255.0f;
  o.z = (float)((c >> 16) & 255u) / 255.0f;
  return o;
}

/* atomic minimum on a 32-bit word; returns the previous value */
int atomicMinInt(int *address, int value);

}  // namespace piko

#endif
