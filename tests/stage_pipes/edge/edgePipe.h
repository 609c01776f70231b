// A pipe for the failure and edge cases of compiled stage code: screen-space
// triangles, a vertex stage that passes them on, and a raster stage that
// fills each triangle's bounding box within its bin.  EDGE_BAD_BIN makes the
// raster stage assign every box from id 7 on to a bin past the grid as well
// (box 7 itself may lie off the screen and never reach it), which the compiled
// schedule must report (the frame returns PIKO_ERR_INVALID) and not follow.
#ifndef TEST_EDGE_PIPE_H
#define TEST_EDGE_PIPE_H

#include "piko/pipe.h"

struct edge_tri : public Primitive {
  int id;
  cvec3f vertPos[3];
  uint32_t color;
};

struct edge_box : public Primitive {
  int id, x0, y0, x1, y1;
  float z;
  uint32_t color;
};

class PassStage : public Stage<32, 32, 64, edge_tri, edge_box> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(edge_tri p) { this->assignToBin(p, 0); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(edge_tri p) {
    edge_box b;
    b.id = p.id;
    b.x0 = (int)piko::fmin2(piko::fmin2(p.vertPos[0].x, p.vertPos[1].x), p.vertPos[2].x);
    b.y0 = (int)piko::fmin2(piko::fmin2(p.vertPos[0].y, p.vertPos[1].y), p.vertPos[2].y);
    b.x1 = (int)piko::fmax2(piko::fmax2(p.vertPos[0].x, p.vertPos[1].x), p.vertPos[2].x) + 1;
    b.y1 = (int)piko::fmax2(piko::fmax2(p.vertPos[0].y, p.vertPos[1].y), p.vertPos[2].y) + 1;
    b.z = p.vertPos[0].z;
    b.color = p.color;
    if (b.x1 > 0 && b.y1 > 0 && b.x0 < constState.screenSizeX && b.y0 < constState.screenSizeY) this->emit(b, 0);
  }
#endif
};

class BoxStage : public Stage<16, 16, 64, edge_box, Pixel> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(edge_box b) {
    const int bx0 = piko::imax(b.x0, 0) / 16, by0 = piko::imax(b.y0, 0) / 16;
    const int bx1 = piko::imin(b.x1 - 1, constState.screenSizeX - 1) / 16;
    const int by1 = piko::imin(b.y1 - 1, constState.screenSizeY - 1) / 16;
    for (int by = by0; by <= by1; ++by)
      for (int bx = bx0; bx <= bx1; ++bx) this->assignToBin(b, by * getNumBinsX() + bx);
#ifdef EDGE_BAD_BIN
    if (b.id >= 7) this->assignToBin(b, getNumBins());
#endif
  }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(edge_box b) {
    const int bx = getBinID() % getNumBinsX(), by = getBinID() / getNumBinsX();
    const int x0 = piko::imax(b.x0, bx * 16), x1 = piko::imin(piko::imin(b.x1, bx * 16 + 16), constState.screenSizeX);
    const int y0 = piko::imax(b.y0, by * 16), y1 = piko::imin(piko::imin(b.y1, by * 16 + 16), constState.screenSizeY);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        int *stored = (int *)&mutableState->zBuffer[y * constState.screenSizeX + x];
        if (piko::atomicMinInt(stored, piko::floatBits(b.z)) < piko::floatBits(b.z)) continue;
        Pixel out;
        out.pos.x = x;
        out.pos.y = y;
        out.color = b.color;
        this->emit(out, 0);
      }
  }
#endif
};

#include "__pikoDefines.h"

class EdgePipe : public PikoPipe {
  PassStage pass;
  BoxStage box;

 public:
  ConstantState *constState_;
  MutableState *mutableState_;
  int count_;
  PikoScreen pikoScreen;
  PikoArray<edge_tri> h_input;

#ifdef __PIKO_DEVICE_MEMBERS__
  __PIKO_DEVICE_MEMBERS__
#endif

  EdgePipe() { pikoConnect(pass, box, 0, 0); }

  void run(ConstantState &constState, MutableState &mutableState, edge_tri *input, int count);
  void allocate(ConstantState &constState, MutableState &mutableState, edge_tri *input, int count);
  void prepare();
  void run_single();
  void destroy();
};

#include "__pikoCompiledPipe.h"

#endif
