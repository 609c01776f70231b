// Host program of the per-pixel-lit raster sample (the raster samples'
// raster_host.h over this sample's pipe); world-space input:
//   raster_lit_sample --scene heightfield [--runs N] [--dump DIR]
#include "rasterPipe.h"
#include "../raster/raster_host.h"

int main(int argc, char **argv) { return raster_host_main<RasterPipe>(argc, argv); }
