/*
 * shared_elementary.h -- force-included into wamp_oracle.c for liborc_wm.so
 * (test infrastructure only): the oracle's pow and tanh become the solver's
 * own (wave-multiphase_amd/csrc/wamp_math.h, built from IEEE operations), so
 * the oracle and the FP64 HIP step evaluate every elementary function alike
 * and a long run can be compared bit for bit.  liborc.so keeps the C
 * library's pow and tanh.
 */
#ifndef ORC_SHARED_ELEMENTARY_H
#define ORC_SHARED_ELEMENTARY_H
#include "../wave-multiphase_amd/csrc/wamp_math.h"
#define pow wm_pow
#define tanh wm_tanh
#endif
