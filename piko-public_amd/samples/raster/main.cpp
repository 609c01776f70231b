// Host program of the raster sample (raster_host.h) over its pipe.
#include "rasterPipe.h"
#include "raster_host.h"

int main(int argc, char **argv) { return raster_host_main<RasterPipe>(argc, argv); }
