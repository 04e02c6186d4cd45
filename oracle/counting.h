// counting.h -- the oracle's counting build (test and measurement
// infrastructure, never the product path).  Included by oracle.cpp when it
// is compiled with -DORACLE_COUNT (oracle/_build/liboracle_count.so): every
// `float` of the restatement then becomes a CountF, whose arithmetic adds to
// the executing thread's counters under the category in effect:
//   + - * / and sqrt        1 FLOP each (no FMA: the restatement is built
//                           without contraction, so a*b+c counts 2)
//   exp log pow sin cos tan acos asin atan atan2
//                           1 transcendental each
//   comparisons, min/max, floor, fabs, negation
//                           not FLOPs (counted as other operations, except
//                           fabs and negation, which are source modifiers on
//                           gfx9)
// The categories split the work of a path: SDF evaluations of the sphere
// tracer (march), central-difference normals, the heterogeneous medium's
// tracking, and everything else (camera, RNG, BSDFs, lights, MIS, colour,
// accumulation).  A fifth, init, holds per-scene constants the restatement
// recomputes where the scene code computes them once in INIT(); it is not
// part of a path's work.  Every other operation counts, constant
// subexpressions included.
#pragma once
#include <cmath>

namespace orccount {
enum { CAT_SHADE, CAT_MARCH, CAT_NORMAL, CAT_MEDIUM, CAT_INIT, CAT_COUNT };
struct Counters {
  long long flops[CAT_COUNT], trans[CAT_COUNT], ops[CAT_COUNT];
  long long steps, normals, paths;
};
extern thread_local Counters tc;
extern thread_local int cat;
struct Scope {
  int prev;
  explicit Scope(int c) : prev(cat) { cat = c; }
  ~Scope() { cat = prev; }
};
inline void flop() { ++tc.flops[cat]; }
inline void trans() { ++tc.trans[cat]; }
inline void op() { ++tc.ops[cat]; }
}  // namespace orccount

struct CountF {
  float v;
  CountF() = default;
  constexpr CountF(float x) : v(x) {}
  constexpr CountF(double x) : v((float)x) {}
  constexpr CountF(int x) : v((float)x) {}
  constexpr CountF(long x) : v((float)x) {}
  constexpr CountF(long long x) : v((float)x) {}
  constexpr CountF(unsigned x) : v((float)x) {}
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  explicit operator double() const { return (double)v; }
  CountF& operator+=(CountF o) { orccount::flop(); v += o.v; return *this; }
  CountF& operator-=(CountF o) { orccount::flop(); v -= o.v; return *this; }
  CountF& operator*=(CountF o) { orccount::flop(); v *= o.v; return *this; }
  CountF& operator/=(CountF o) { orccount::flop(); v /= o.v; return *this; }
};
static_assert(sizeof(CountF) == sizeof(float), "CountF must be laid out as a float");
inline CountF operator+(CountF a, CountF b) { orccount::flop(); return CountF(a.v + b.v); }
inline CountF operator-(CountF a, CountF b) { orccount::flop(); return CountF(a.v - b.v); }
inline CountF operator*(CountF a, CountF b) { orccount::flop(); return CountF(a.v * b.v); }
inline CountF operator/(CountF a, CountF b) { orccount::flop(); return CountF(a.v / b.v); }
inline CountF operator-(CountF a) { return CountF(-a.v); }
inline bool operator<(CountF a, CountF b) { orccount::op(); return a.v < b.v; }
inline bool operator>(CountF a, CountF b) { orccount::op(); return a.v > b.v; }
inline bool operator<=(CountF a, CountF b) { orccount::op(); return a.v <= b.v; }
inline bool operator>=(CountF a, CountF b) { orccount::op(); return a.v >= b.v; }
inline bool operator==(CountF a, CountF b) { orccount::op(); return a.v == b.v; }
inline bool operator!=(CountF a, CountF b) { orccount::op(); return a.v != b.v; }

namespace std {
inline CountF sqrt(CountF a) { orccount::flop(); return CountF(std::sqrt(a.v)); }
inline CountF exp(CountF a) { orccount::trans(); return CountF(std::exp(a.v)); }
inline CountF log(CountF a) { orccount::trans(); return CountF(std::log(a.v)); }
inline CountF pow(CountF a, CountF b) { orccount::trans(); return CountF(std::pow(a.v, b.v)); }
inline CountF sin(CountF a) { orccount::trans(); return CountF(std::sin(a.v)); }
inline CountF cos(CountF a) { orccount::trans(); return CountF(std::cos(a.v)); }
inline CountF tan(CountF a) { orccount::trans(); return CountF(std::tan(a.v)); }
inline CountF acos(CountF a) { orccount::trans(); return CountF(std::acos(a.v)); }
inline CountF asin(CountF a) { orccount::trans(); return CountF(std::asin(a.v)); }
inline CountF atan(CountF a) { orccount::trans(); return CountF(std::atan(a.v)); }
inline CountF atan2(CountF a, CountF b) { orccount::trans(); return CountF(std::atan2(a.v, b.v)); }
inline CountF floor(CountF a) { orccount::op(); return CountF(std::floor(a.v)); }
inline CountF fabs(CountF a) { return CountF(std::fabs(a.v)); }
inline CountF fmax(CountF a, CountF b) { orccount::op(); return CountF(std::fmax(a.v, b.v)); }
inline CountF fmin(CountF a, CountF b) { orccount::op(); return CountF(std::fmin(a.v, b.v)); }
}  // namespace std

#define float CountF
