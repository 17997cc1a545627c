// Per-kernel constants of the packed record, shared by the posterior kernel
// and the record packer of the differentiable path.
#pragma once

#include "common.hpp"

namespace lakde {

// digamma for x > 0: upward recurrence to x >= 6, then the asymptotic series
// psi(x) = ln x - 1/(2x) - sum_k B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18),
// truncated after the x^-10 term (error < 1e-11 for x >= 6).
__device__ inline double digamma_d(double x) {
  double r = 0.0;
  while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
  double f = 1.0 / (x * x);
  double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132)))));
  return r + log(x) - 0.5 / x + t;
}

// row i and column r of entry t of a row-packed lower triangle
// (t = i (i + 1) / 2 + r, r <= i); the float estimate is corrected exactly
__device__ __forceinline__ void tri_index(int t, int *i, int *r) {
  int k = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
  if ((k + 1) * (k + 2) / 2 <= t) ++k;
  if (k * (k + 1) / 2 > t) --k;
  *i = k;
  *r = t - k * (k + 1) / 2;
}

// cst[0..3] = E-step const, E-step exponent, predictive coefficient,
// predictive exponent, in the float32 operation order of
// lakde/local_adaptive_kde.py:36-44 (log rho) and :209-217 (predictive).
__device__ inline void record_consts(float *cst, float nu, float logdet, int D) {
  const float half_nu = nu / 2.f;
  double mvd = 0.0;
  for (int i = 1; i <= D; ++i) mvd += digamma_d((double)(half_nu + (1.f - (float)i) / 2.f));
  const float eld = ((float)mvd + (float)(D * 0.69314718055994530942)) + logdet;
  const float h1 = (nu + 1.f) / 2.f;
  const float h2 = ((nu + 1.f) - (float)D) / 2.f;
  const float lg1 = (float)lgamma((double)h1);
  const float lg2 = (float)lgamma((double)h2);
  cst[0] = 0.5f * eld;
  cst[1] = -0.5f * nu;
  cst[2] = ((lg1 - lg2) + 0.5f * logdet) - (float)(D / 2.0 * 1.14472988584940017414);
  cst[3] = h1;
}

// record_consts with the D digamma terms and the two lgamma calls spread
// over the lanes of a wave (every lane holds the same nu and logdet; D <= 62);
// the terms are summed in the same order, so the constants are identical.
// Lane 0 writes cst.
__device__ inline void record_consts_wave(float *cst, float nu, float logdet, int D, int lane) {
  const float half_nu = nu / 2.f;
  const float h1 = (nu + 1.f) / 2.f;
  const float h2 = ((nu + 1.f) - (float)D) / 2.f;
  double t = 0.0;
  if (lane < D) t = digamma_d((double)(half_nu + (1.f - (float)(lane + 1)) / 2.f));
  else if (lane <= D + 1) t = lgamma((double)(lane == D ? h1 : h2));
  double mvd = 0.0;
  for (int i = 0; i < D; ++i) mvd += __shfl(t, i, kWave);
  const float lg1 = (float)__shfl(t, D, kWave);
  const float lg2 = (float)__shfl(t, D + 1, kWave);
  if (lane == 0) {
    const float eld = ((float)mvd + (float)(D * 0.69314718055994530942)) + logdet;
    cst[0] = 0.5f * eld;
    cst[1] = -0.5f * nu;
    cst[2] = ((lg1 - lg2) + 0.5f * logdet) - (float)(D / 2.0 * 1.14472988584940017414);
    cst[3] = h1;
  }
}

}  // namespace lakde
