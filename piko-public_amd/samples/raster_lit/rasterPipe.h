// The per-pixel-lit raster sample: VertexShader -> Raster -> PikoScreen, the
// raster sample's graph with stage bodies of its own -- the vertex stage
// passes the normals on and the raster stage lights every pixel from the
// interpolated normal (what the reference's raster stage does when GORAUD
// is not defined, raster.pikostage:186-191, 294-304).  The kernel library
// implements only Gouraud shading, so pikoc_hip compiles these bodies.
#ifndef SAMPLE_RASTER_LIT_PIPE_H
#define SAMPLE_RASTER_LIT_PIPE_H

#include "piko/pipe.h"

#define PIKO_SUBPIXEL_BITS 8

#include "rasterTypes.h"
#include "vertexShader.h"
#include "raster.h"

#include "__pikoDefines.h"

class RasterPipe : public PikoPipe {
  VertexShaderStage vertexShader;
  RasterStage raster;

 public:
  ConstantState *constState_;
  MutableState *mutableState_;
  int count_;
  PikoScreen pikoScreen;
  PikoArray<raster_wtri> h_input;

#ifdef __PIKO_DEVICE_MEMBERS__
  __PIKO_DEVICE_MEMBERS__
#endif

  RasterPipe() { pikoConnect(vertexShader, raster, 0, 0); }

  void run(ConstantState &constState, MutableState &mutableState, raster_wtri *input, int count);
  void allocate(ConstantState &constState, MutableState &mutableState, raster_wtri *input, int count);
  void prepare();
  void run_single();
  void destroy();
};

#include "__pikoCompiledPipe.h"

#endif
