/*
 * opcount.cpp -- counters of the operation-counting oracle build
 * (test infrastructure only; see opcount.h).
 */
#include "opcount.h"

uint64_t orc_opcount_[OPC_N];

extern "C" void orc_opcount_read(uint64_t* out) { memcpy(out, orc_opcount_, sizeof(orc_opcount_)); }
extern "C" void orc_opcount_reset(void) { memset(orc_opcount_, 0, sizeof(orc_opcount_)); }
extern "C" int orc_opcount_categories(void) { return OPC_N; }
