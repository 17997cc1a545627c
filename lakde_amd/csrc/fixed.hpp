// Order-independent column sums.
//
// The per-column sums that feed the normaliser (the lazy rebalance's true
// and stored column sums, the block's old column sums, the fold totals) are
// accumulated by many threads at once.  Floating-point atomics would make
// them depend on the order the adds arrive in, and so make a fit differ from
// run to run.  Here every value (a non-negative float) is converted exactly
// to a 128-bit fixed-point integer with 2^-80 resolution and added with two
// 64-bit integer atomics (the low word, then the high word plus the low
// word's carry).  Integer addition is associative and commutative, so the
// total is the same whatever the order; a value added once and removed later
// cancels exactly.  Floats below 2^-104 lose bits below the resolution
// (deterministically); values must stay below 2^46 (the callers' values are
// true responsibilities <= 1 and stored values <= 2^40, the lazy
// rebalance's unsafe-step window), and NaN, infinities, negative and larger
// values set the top bit, which reads back as NaN.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lakde_amd.h"

namespace lakde {

typedef lakde_fx128 Fx;  // { uint64_t lo, hi }

constexpr unsigned long long kFxFlag = 1ull << 63;

// v * 2^80 as a 128-bit integer (hi:lo); *bad for values it cannot hold
__device__ __forceinline__ void fx_from_float(float v, unsigned long long *lo,
                                              unsigned long long *hi, bool *bad) {
  *lo = 0ull;
  *hi = 0ull;
  *bad = !(v >= 0.f && v < 0x1p46f);  // also NaN
  if (*bad || v == 0.f) return;
  const uint32_t bits = __float_as_uint(v);
  const int ex = (int)((bits >> 23) & 0xffu);
  unsigned long long mant = bits & 0x7fffffu;
  int e2;  // v = mant * 2^e2
  if (ex == 0) {
    e2 = -149;
  } else {
    mant |= 0x800000u;
    e2 = ex - 150;
  }
  const int sh = e2 + 80;  // v * 2^80 = mant << sh, with mant < 2^24 and sh <= 101
  if (sh < 0) {
    *lo = sh <= -64 ? 0ull : (mant >> (-sh));
  } else if (sh < 64) {
    *lo = mant << sh;
    *hi = sh > 40 ? (mant >> (64 - sh)) : 0ull;
  } else {
    *hi = mant << (sh - 64);
  }
}

__device__ __forceinline__ void fx_atomic_add(Fx *acc, float v) {
  unsigned long long lo, hi;
  bool bad;
  fx_from_float(v, &lo, &hi, &bad);
  unsigned long long *w = reinterpret_cast<unsigned long long *>(acc);
  if (bad) {
    atomicOr(&w[1], kFxFlag);
    return;
  }
  if (lo != 0ull) {
    const unsigned long long old = atomicAdd(&w[0], lo);
    if (old + lo < old) hi += 1ull;
  }
  if (hi != 0ull) atomicAdd(&w[1], hi);
}

// the exact sum a - b (b a sub-sum of a) rounded to double; NaN if either
// was flagged
__device__ __forceinline__ double fx_diff(Fx a, Fx b) {
  if (((a.hi | b.hi) & kFxFlag) != 0ull) return __builtin_nan("");
  const unsigned long long lo = a.lo - b.lo;
  const unsigned long long borrow = a.lo < b.lo ? 1ull : 0ull;
  const long long hi = (long long)(a.hi - b.hi - borrow);
  return (double)hi * 0x1p-16 + (double)lo * 0x1p-80;
}

__device__ __forceinline__ double fx_value(Fx a) {
  if ((a.hi & kFxFlag) != 0ull) return __builtin_nan("");
  return (double)a.hi * 0x1p-16 + (double)a.lo * 0x1p-80;
}

__device__ __forceinline__ Fx fx_zero() {
  Fx z;
  z.lo = 0;
  z.hi = 0;
  return z;
}

// a - b + c exactly (the flag of any operand kept)
__device__ __forceinline__ Fx fx_sub_add(Fx a, Fx b, Fx c) {
  Fx r;
  const unsigned long long flag = (a.hi | b.hi | c.hi) & kFxFlag;
  unsigned long long lo = a.lo - b.lo;
  unsigned long long hi = a.hi - b.hi - (a.lo < b.lo ? 1ull : 0ull);
  const unsigned long long lo2 = lo + c.lo;
  hi += c.hi + (lo2 < lo ? 1ull : 0ull);
  r.lo = lo2;
  r.hi = (hi & ~kFxFlag) | flag;
  return r;
}

}  // namespace lakde
