/*
 * opcount.h -- operation counting for the CPU ORACLE (test infrastructure only).
 *
 * Force-included (g++ -include) in front of wamp_oracle.c to build
 * liborc_count.so: every `double` of the oracle becomes `CountedD`, a
 * one-double struct whose operators and math functions do the same IEEE
 * operation as the plain build and bump a per-category counter.  The
 * counting build therefore computes bit for bit what liborc.so computes
 * (checked by a test) and reports how many floating-point operations of each
 * kind the reference's unfused step performs.  Built without OpenMP: the
 * counters are plain globals and the loops run serially.
 */
#ifndef WAMP_ORACLE_OPCOUNT_H
#define WAMP_ORACLE_OPCOUNT_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#include <omp.h>
#include <cmath>

enum {
  OPC_ADD, OPC_MUL, OPC_DIV, OPC_SQRT, OPC_POW, OPC_TANH, OPC_FABS, OPC_FLOOR, OPC_CMP, OPC_N
};
extern uint64_t orc_opcount_[OPC_N];

struct CountedD {
  double v;
  CountedD() = default;
  constexpr CountedD(double x) : v(x) {}
  explicit operator double() const { return v; }
  CountedD& operator+=(CountedD o) { ++orc_opcount_[OPC_ADD]; v = v + o.v; return *this; }
  CountedD& operator-=(CountedD o) { ++orc_opcount_[OPC_ADD]; v = v - o.v; return *this; }
  CountedD& operator*=(CountedD o) { ++orc_opcount_[OPC_MUL]; v = v * o.v; return *this; }
  CountedD& operator/=(CountedD o) { ++orc_opcount_[OPC_DIV]; v = v / o.v; return *this; }
};
static_assert(sizeof(CountedD) == sizeof(double), "CountedD must have the layout of double");

inline CountedD operator+(CountedD a, CountedD b) { ++orc_opcount_[OPC_ADD]; return a.v + b.v; }
inline CountedD operator-(CountedD a, CountedD b) { ++orc_opcount_[OPC_ADD]; return a.v - b.v; }
inline CountedD operator*(CountedD a, CountedD b) { ++orc_opcount_[OPC_MUL]; return a.v * b.v; }
inline CountedD operator/(CountedD a, CountedD b) { ++orc_opcount_[OPC_DIV]; return a.v / b.v; }
inline CountedD operator-(CountedD a) { return -a.v; }
inline CountedD operator+(CountedD a) { return a; }
inline bool operator<(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v < b.v; }
inline bool operator>(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v > b.v; }
inline bool operator<=(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v <= b.v; }
inline bool operator>=(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v >= b.v; }
inline bool operator==(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v == b.v; }
inline bool operator!=(CountedD a, CountedD b) { ++orc_opcount_[OPC_CMP]; return a.v != b.v; }

inline CountedD sqrt(CountedD a) { ++orc_opcount_[OPC_SQRT]; return std::sqrt(a.v); }
inline CountedD pow(CountedD a, CountedD b) { ++orc_opcount_[OPC_POW]; return std::pow(a.v, b.v); }
inline CountedD tanh(CountedD a) { ++orc_opcount_[OPC_TANH]; return std::tanh(a.v); }
inline CountedD fabs(CountedD a) { ++orc_opcount_[OPC_FABS]; return std::fabs(a.v); }
inline CountedD floor(CountedD a) { ++orc_opcount_[OPC_FLOOR]; return std::floor(a.v); }
inline bool isnan(CountedD a) { ++orc_opcount_[OPC_CMP]; return std::isnan(a.v); }

#ifdef __cplusplus
extern "C" {
#endif
/* Copy the counters (OPC_N of them, in the enum's order) to out. */
void orc_opcount_read(uint64_t* out);
void orc_opcount_reset(void);
int orc_opcount_categories(void);
#ifdef __cplusplus
}
#endif

#define double CountedD

#endif
