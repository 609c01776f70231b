// Primitive types of the Reyes sample (reyesPipe/pikoTypes.h:65-79 of the
// reference): the bicubic patch and the micropolygon, plus the shaded
// sample the Shade stage emits.
#ifndef SAMPLE_REYES_TYPES_H
#define SAMPLE_REYES_TYPES_H

#include "piko/piko.h"

struct piko_patch : public Primitive {
  cvec4f CP[16];
  int numSplits;
  int nu;
  int nv;
  int id;
  cvec2f bbmin;
  cvec2f bbmax;
};

struct piko_upoly : public Primitive {
  cvec4f screenPos[4];
  cvec3f normal[4];
};

// One covered sample that passed the depth test: the PikoScreen keeps
// PIKO_REYES_SPP float RGB samples per pixel and averages them (a box
// filter) into the frame's FP32 colour.
struct piko_sample : public Primitive {
  cvec2i pos;
  int index;
  float rgb[3];
};

#endif
