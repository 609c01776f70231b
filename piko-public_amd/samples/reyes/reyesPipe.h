// The Reyes sample pipe: Split (loop) -> Dice -> Shade -> PikoScreen (the
// pipe class shape of reyesPipe/reyesPipe.h:12-40), with the settings of the
// 4,096-patch sphere workload: split below 16 px, 8x8 dice, 4 samples per
// pixel, displacement 0.05.
#ifndef SAMPLE_REYES_PIPE_H
#define SAMPLE_REYES_PIPE_H

#include "piko/pipe.h"

#define THRESHOLD_PIXELS 16.0f
#define PIKO_REYES_DICE 8
#define PIKO_REYES_SPP 4
#define PIKO_REYES_DISPLACEMENT 0.05f

#include "reyesTypes.h"
#include "stages.h"

#include "__pikoDefines.h"

class ReyesPipe : public PikoPipe {
  SplitStage split;
  DiceStage dice;
  ShadeStage shade;

 public:
  ConstantState *constState_;
  MutableState *mutableState_;
  int count_;
  PikoScreen pikoScreen;
  PikoArray<piko_patch> h_input;

#ifdef __PIKO_DEVICE_MEMBERS__
  __PIKO_DEVICE_MEMBERS__
#endif

  ReyesPipe() {
    pikoConnect(split, dice, 0, 0);
    pikoConnect(dice, shade, 0, 0);
    pikoConnect(split, split, 1, 1);
  }

  void run(ConstantState &constState, MutableState &mutableState, piko_patch *input, int count);
  void allocate(ConstantState &constState, MutableState &mutableState, piko_patch *input, int count);
  void prepare();
  void run_single();
  void destroy();
};

#include "__pikoCompiledPipe.h"

#endif
