// Stages of the Reyes sample: Split (looping on itself through out port 1),
// Dice, Shade.
#ifndef SAMPLE_REYES_STAGES_H
#define SAMPLE_REYES_STAGES_H

#include "reyesTypes.h"

#define SPLIT_BINSIZE 0
#define DICE_BINSIZE 64
#define SHADE_BINSIZE 64

class SplitStage : public Stage<SPLIT_BINSIZE, SPLIT_BINSIZE, 256, piko_patch, piko_patch> {
 public:
  void assignBin(piko_patch p) { this->assignToBin(p, 0); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(piko_patch p) {}
};

class DiceStage : public Stage<DICE_BINSIZE, DICE_BINSIZE, 128, piko_patch, piko_upoly> {
 public:
  void assignBin(piko_patch p) { this->assignToBin(p, (p.id / 128)); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(piko_patch p) {}
};

class ShadeStage : public Stage<SHADE_BINSIZE, SHADE_BINSIZE, 512, piko_upoly, Pixel> {
 public:
  void assignBin(piko_upoly p) { this->assignToBin(p, 0); }
  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }
  void process(piko_upoly p) {}
};

#endif
