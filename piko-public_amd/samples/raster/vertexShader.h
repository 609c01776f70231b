// Vertex stage of the raster sample: per-triangle transform, culls and
// fixed-point setup; triangles go to chunked bins of 512 consecutive ids.
#ifndef SAMPLE_VERTEX_SHADER_H
#define SAMPLE_VERTEX_SHADER_H

#include "rasterTypes.h"

#define VS_BINSIZE 32

class VertexShaderStage : public Stage<VS_BINSIZE, VS_BINSIZE, 512, raster_wtri, raster_stri> {
 public:
  void assignBin(raster_wtri p) { this->assignToBin(p, p.id / 512); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(raster_wtri p) {}
};

#endif
