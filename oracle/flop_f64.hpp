/*
 * TEST INFRASTRUCTURE ONLY -- an FP64 number that counts its arithmetic.
 *
 * oracle/flopcount.cpp compiles the oracle's C sources as C++ with `double`
 * defined as F64, so every +, -, *, / the restatement performs (in the
 * reference's operation order) is counted once, and sqrt, log, pow, exp
 * are counted apart.  Comparisons, negation, fabs, fmin and fmax are not
 * arithmetic and are not counted.  Counts are kept per phase (ORC_PHASE in
 * oracle.h marks the phases of the step).
 */
#pragma once
#include <math.h>
#include <cmath>

namespace flop {
enum { ADD, MUL, DIV, SQRT, LOG, POW, EXP, NKIND };
constexpr int NPHASE = 8;
extern long long count[NPHASE][NKIND];
extern int phase;
inline void tick(int k) { count[phase][k] += 1; }
}  // namespace flop

struct F64 {
  double v;
  F64() = default;
  F64(double x) : v(x) {}
  explicit operator double() const { return v; }
  F64& operator+=(F64 o) { flop::tick(flop::ADD); v = v + o.v; return *this; }
  F64& operator-=(F64 o) { flop::tick(flop::ADD); v = v - o.v; return *this; }
  F64& operator*=(F64 o) { flop::tick(flop::MUL); v = v * o.v; return *this; }
  F64& operator/=(F64 o) { flop::tick(flop::DIV); v = v / o.v; return *this; }
};
inline F64 operator+(F64 a, F64 b) { flop::tick(flop::ADD); return F64(a.v + b.v); }
inline F64 operator-(F64 a, F64 b) { flop::tick(flop::ADD); return F64(a.v - b.v); }
inline F64 operator*(F64 a, F64 b) { flop::tick(flop::MUL); return F64(a.v * b.v); }
inline F64 operator/(F64 a, F64 b) { flop::tick(flop::DIV); return F64(a.v / b.v); }
inline F64 operator-(F64 a) { return F64(-a.v); }
inline bool operator<(F64 a, F64 b) { return a.v < b.v; }
inline bool operator>(F64 a, F64 b) { return a.v > b.v; }
inline bool operator<=(F64 a, F64 b) { return a.v <= b.v; }
inline bool operator>=(F64 a, F64 b) { return a.v >= b.v; }
inline bool operator==(F64 a, F64 b) { return a.v == b.v; }
inline bool operator!=(F64 a, F64 b) { return a.v != b.v; }
inline F64 sqrt(F64 a) { flop::tick(flop::SQRT); return F64(::sqrt(a.v)); }
inline F64 log(F64 a) { flop::tick(flop::LOG); return F64(::log(a.v)); }
inline F64 exp(F64 a) { flop::tick(flop::EXP); return F64(::exp(a.v)); }
inline F64 pow(F64 a, F64 b) { flop::tick(flop::POW); return F64(::pow(a.v, b.v)); }
inline F64 fabs(F64 a) { return F64(::fabs(a.v)); }
inline F64 fmin(F64 a, F64 b) { return F64(::fmin(a.v, b.v)); }
inline F64 fmax(F64 a, F64 b) { return F64(::fmax(a.v, b.v)); }
inline bool isfinite(F64 a) { return std::isfinite(a.v); }
inline bool isnan(F64 a) { return std::isnan(a.v); }
