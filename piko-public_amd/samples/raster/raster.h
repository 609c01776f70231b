// Raster stage of the raster sample: trivial-reject bin walk in
// assignBin, coverage, depth test and Gouraud colour in process.
#ifndef SAMPLE_RASTER_H
#define SAMPLE_RASTER_H

#include "rasterTypes.h"

#define RASTER_BINSIZE_LG2 3
#define RASTER_BINSIZE (1 << RASTER_BINSIZE_LG2)
#define RASTER_THREADCOUNT 64

class RasterStage : public Stage<RASTER_BINSIZE, RASTER_BINSIZE, RASTER_THREADCOUNT, raster_stri, Pixel> {
 public:
  void assignBin(raster_stri p) { this->assignToBin(p, 0); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(raster_stri p) {}
};

#endif
