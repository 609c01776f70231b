// Primitive types of the per-pixel-lit raster sample: the world-space input
// triangle (the raster sample's) and a setup triangle that carries the
// vertex normals to the raster stage in place of lit vertex colours.
#ifndef SAMPLE_RASTER_LIT_TYPES_H
#define SAMPLE_RASTER_LIT_TYPES_H

#include "piko/piko.h"

struct raster_wtri : public Primitive {
  int id;
  cvec3f vertPos[3];     // world space
  cvec3f vertNormal[3];  // world space
};

struct raster_stri : public Primitive {
  int32_t xy[6];  // fixed point, PIKO_SUBPIXEL_BITS fractional bits
  float z[3];
  cvec3f nrm[3];  // the vertex normals, unlit
};

#endif
