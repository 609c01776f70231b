// The raster sample pipe: VertexShader -> Raster -> PikoScreen
// (the pipe class shape of rasterPipelineFixPt/rasterPipe.h:16-45).
#ifndef SAMPLE_RASTER_PIPE_H
#define SAMPLE_RASTER_PIPE_H

#include "piko/pipe.h"

#define GORAUD
#ifndef RASTER_WORLD_INPUT
#define VTX_PRETRANSFORM
#endif
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
