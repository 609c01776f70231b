// Primitive types of the Reyes sample (reyesPipe/pikoTypes.h:65-79 of the
// reference): the bicubic patch and the micropolygon.
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

#endif
