/*
 * piko/stage.h -- Stage, the directives of the assignBin / schedule phases,
 * PikoScreen and PikoArray (api/include/piko/stage.h:53-391,
 * api/include/piko/deviceFunctions.h:22-50, api/include/internal/
 * datatypes.h:114-144).
 *
 * pikoc_hip reads a stage class's bin size, threads per tile, primitive
 * types, directives and bodies from the source (with __PIKOC_DEVICE__
 * defined).  It maps the pipe onto the kernel library when every stage body
 * is one the library implements, and otherwise compiles the bodies for
 * gfx950, where piko/stage_exec.h defines the members declared below under
 * __PIKOC_DEVICE__.  The directive functions do nothing: the schedule is
 * the library's or stage_exec.h's.
 */
#ifndef PIKO_STAGE_H
#define PIKO_STAGE_H

#include <string.h>

#include <vector>

#include "piko/builtinTypes.h"
#ifdef __PIKOC_DEVICE__
#include "piko/device.h"
#endif

enum AssignPolicy { PREVIOUS_BINS, BOUNDING_BOX, POSITION };
enum WaitPolicy { CUSTOM, BATCH, END_STAGE, END_BIN };
enum SchedulePolicy { LOAD_BALANCE, ROUND_ROBIN, SERIAL, ALL };

struct StageBase {
  StageBase *outPort[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  StageBase *inPort[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

template <int BinSizeX, int BinSizeY, int ThreadsPerTile, class InPrim, class OutPrim>
class Stage : public StageBase {
 public:
  typedef InPrim input_type;
  typedef OutPrim output_type;
  static constexpr int binSizeX = BinSizeX, binSizeY = BinSizeY, threadsPerTile = ThreadsPerTile;

 protected:
  void specifySchedule(SchedulePolicy, int tileSplitSize = 0) { (void)tileSplitSize; }
  void specifyWait(WaitPolicy) {}
  void specifyMaxOutPrims(int) {}
#ifndef __PIKOC_DEVICE__
  void assignToBin(const InPrim &, int) {}
  void assignToBin(const InPrim &, AssignPolicy) {}
#else
  /* What a stage body sees (stage.h:22-24, deviceFunctions.h:63-109): bin
   * assignment, the frame's uniforms, the mutable state (the depth buffer),
   * the bin being processed and the bin grid, and emit to an output port.
   * Declared only: whatever executes the bodies defines them (the compiled
   * stage code in piko/stage_exec.h, the host harness of the tests). */
  void assignToBin(const InPrim &p, int bin);
  void assignToBin(const InPrim &p, AssignPolicy policy);
  ConstantState constState;
  MutableState *mutableState;
  void *pikoRun_ = nullptr;  // the executing schedule's state for this body
  int getBinID() const;
  int getNumBins() const;
  int getNumBinsX() const;
  int getNumBinsY() const;
  void emit(const OutPrim &p, int outPort);
#endif
};

/* PikoScreen (stage.h:53-136): the colour target.  getData() returns the last
 * frame's colour words (ABGR8, row 0 at the bottom), owned by the screen. */
class PikoScreen {
 public:
  unsigned *getData() {
    if (fetch_) fetch_(ctx_, data_);
    return data_.empty() ? nullptr : data_.data();
  }
  /* Bound by the generated allocate(). */
  void bind(void *ctx, int (*fetch)(void *, std::vector<unsigned> &), size_t pixels) {
    ctx_ = ctx;
    fetch_ = fetch;
    data_.assign(pixels, 0u);
  }

 private:
  void *ctx_ = nullptr;
  int (*fetch_)(void *, std::vector<unsigned> &) = nullptr;
  std::vector<unsigned> data_;
};

/* PikoArray (datatypes.h:114-144): the pipe's host-side input.  copyData keeps
 * every primitive (the reference truncates at MAX_NUM_PRIMS without a word). */
template <typename T>
class PikoArray {
 public:
  void copyData(const T *in, int count) { data_.assign(in, in + count); }
  int count() const { return (int)data_.size(); }
  const T *data() const { return data_.data(); }

 private:
  std::vector<T> data_;
};

#endif
