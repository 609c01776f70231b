/*
 * piko/pipe.h -- PikoPipe and pikoConnect (api/include/piko/pipe.h).
 */
#ifndef PIKO_PIPE_H
#define PIKO_PIPE_H

#include "piko/stage.h"

/* Output port outPortNum of outStg feeds input port inPortNum of inStg. */
template <class S1, class S2>
static void pikoConnect(S1 &outStg, S2 &inStg, const int outPortNum, const int inPortNum) {
  outStg.outPort[outPortNum] = &inStg;
  inStg.inPort[inPortNum] = &outStg;
}

class PikoPipe {};

#endif
