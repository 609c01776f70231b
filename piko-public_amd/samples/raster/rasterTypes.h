// Primitive types of the raster sample (basicTypes/rasterTypes.h:8-71 of the
// reference): the input triangle and the setup triangle the vertex stage
// emits to the raster stage.
#ifndef SAMPLE_RASTER_TYPES_H
#define SAMPLE_RASTER_TYPES_H

#include "piko/piko.h"

struct raster_wtri : public Primitive {
  int id;
  cvec3f vertPos[3];  // screen-space (VTX_PRETRANSFORM) or world-space
#ifdef VTX_PRETRANSFORM
  uint32_t vertColor[3];  // packed ABGR8
#else
  cvec3f vertNormal[3];
#endif
};

struct raster_stri : public Primitive {
  int32_t xy[6];  // fixed point, PIKO_SUBPIXEL_BITS fractional bits
  float z[3];
  uint32_t col[3];
};

#endif
