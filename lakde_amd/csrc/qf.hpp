// Per-pair quadratic form and epilogues shared by the E-step and predictive
// kernels.  Every value that two kernels must agree on bit-for-bit (the
// E-step ballot and the compaction that recomputes the kept entries) is
// produced by these functions with explicit fmaf/__fmul_rn/__fadd_rn, and the
// library is built with -ffp-contract=off, so no context-dependent
// contraction can change a result between call sites.
#pragma once

#include "common.hpp"

namespace lakde {

// Q = || Linv (x_m - x_n) ||^2 with Linv packed row-major lower in rec.
// rec is wave-uniform at every call site, so its loads are scalar loads and
// the factor entries reach the FMAs as SGPR operands.
template <int D>
__device__ __forceinline__ float qf_rec(const float *__restrict__ rec, const float (&xm)[D]) {
  float d[D];
#pragma unroll
  for (int j = 0; j < D; ++j) d[j] = xm[j] - rec[rec_x(D) + j];
  float q = 0.f;
  int t = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    float y = 0.f;
#pragma unroll
    for (int j = 0; j <= i; ++j) y = fmaf(rec[t + j], d[j], y);
    t += i + 1;
    q = fmaf(y, y, q);
  }
  return q;
}

// log rho_nm = Q * (-nu/2) + const  (two roundings, as Q *= e; Q += c)
template <int D>
__device__ __forceinline__ float log_rho(const float *__restrict__ rec, const float (&xm)[D]) {
  const float q = qf_rec<D>(rec, xm);
  return __fadd_rn(__fmul_rn(q, rec[rec_c(D) + 1]), rec[rec_c(D) + 0]);
}

// log rho from a quadratic form already computed (the same two roundings)
template <int D>
__device__ __forceinline__ float log_rho_q(const float *__restrict__ rec, float q) {
  return __fadd_rn(__fmul_rn(q, rec[rec_c(D) + 1]), rec[rec_c(D) + 0]);
}

// log p_n(y_m) = coeff - (nu+1)/2 * log1p(Q)
template <int D>
__device__ __forceinline__ float log_pred(const float *__restrict__ rec, const float (&ym)[D]) {
  const float q = qf_rec<D>(rec, ym);
  return __fsub_rn(rec[rec_c(D) + 2], __fmul_rn(rec[rec_c(D) + 3], log1pf(q)));
}

// responsibility and keep decision (lakde/local_adaptive_kde.py:138-142):
// threshold > 0 keeps r > threshold; threshold == 0 keeps r != 0.
__device__ __forceinline__ bool keep_resp(float d, float threshold, float log_t_lo, float *r) {
  if (threshold > 0.f) {
    if (!(d > log_t_lo)) return false;
    *r = expf(d);
    return *r > threshold;
  }
  // exp(d) rounds to a nonzero float iff exp(d) > 2^-150, i.e. d > -150 ln 2;
  // below ln(2^-149) the correctly rounded result is the smallest denormal
  // (device expf returns 0 there, a denormal ulp away)
  if (d <= -103.972077f) return false;
  *r = (d < -103.278929f) ? 1.40129846e-45f : expf(d);
  return true;
}

// online log-sum-exp state
struct Lse {
  float m, s;
  __device__ __forceinline__ void init() { m = neg_inf(); s = 0.f; }
  // fold a group of 4 values
  __device__ __forceinline__ void add4(float x0, float x1, float x2, float x3) {
    float g = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
    if (g > m) { s = s * __expf(m - g); m = g; }
    if (m != neg_inf()) s += (__expf(x0 - m) + __expf(x1 - m)) + (__expf(x2 - m) + __expf(x3 - m));
  }
  __device__ __forceinline__ void add1(float x) {
    if (x > m) { s = s * __expf(m - x); m = x; }
    if (m != neg_inf()) s += __expf(x - m);
  }
  __device__ __forceinline__ void merge(float m2, float s2) {
    if (m2 == neg_inf()) return;
    if (m2 > m) { s = s * __expf(m - m2) + s2; m = m2; }
    else s += s2 * __expf(m2 - m);
  }
  __device__ __forceinline__ float value() const { return m == neg_inf() ? m : m + logf(s); }
};

// Streaming log-sum-exp of the E-step's hot loop, with the arithmetic of
// Lse::add4: a new maximum is rare, so its rescale is a real branch, and
// the terms subtract a copy mz of the maximum that starts at 0 instead of
// -inf, so that a -inf term (the excluded diagonal) gives exactly 0 without
// a test.
struct LseE {
  float m, mz, s;
  __device__ __forceinline__ void init() { m = neg_inf(); mz = 0.f; s = 0.f; }
  // returns whether the maximum moved
  __device__ __forceinline__ bool fold(float g) {
    if (g > m) {
      asm volatile("");  // keep the rescale a branch instead of speculating it
      s = s * __expf(m - g);
      m = g;
      mz = g;
      return true;
    }
    return false;
  }
  __device__ __forceinline__ float term(float x) const { return __expf(x - mz); }
};

}  // namespace lakde

// All dimensions 1..63 supported by the fused kernels.
#define LAKDE_FOR_EACH_D(X)                                                                   \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)     \
  X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)  \
  X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39) X(40) X(41) X(42) X(43) X(44) X(45) X(46)  \
  X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58) X(59) X(60) X(61)  \
  X(62) X(63)
