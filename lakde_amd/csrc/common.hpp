// Shared host/device helpers for the lakde_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/lakde_amd.h"
#include "fixed.hpp"

namespace lakde {

// ------------------------------------------------------------ host errors

void set_error(const std::string &msg);

#define LAKDE_CHECK_ARG(cond, msg)                                  \
  do {                                                              \
    if (!(cond)) {                                                  \
      ::lakde::set_error(std::string(__func__) + ": " + (msg));     \
      return LAKDE_EINVAL;                                          \
    }                                                               \
  } while (0)

#define LAKDE_HIP(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::lakde::set_error(std::string(__func__) + ": " #call ": " +             \
                         hipGetErrorString(e_));                               \
      return LAKDE_EHIP;                                                       \
    }                                                                          \
  } while (0)

// after a kernel launch
#define LAKDE_LAUNCHED() LAKDE_HIP(hipGetLastError())

inline hipStream_t as_stream(lakde_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------- device helpers

constexpr int kWave = 64;

__device__ __forceinline__ float neg_inf() { return -__builtin_huge_valf(); }

// torch.logaddexp semantics
__device__ __forceinline__ float logaddexp_f(float a, float b) {
  if (a == b && __builtin_isinf(a)) return a;
  float m = fmaxf(a, b);
  float d = -fabsf(a - b);
  return m + log1pf(expf(d));
}

// a + log1p(clamp_min(-exp(b - a), -1)), NaN propagating like torch.clamp_min
// (lakde/linalg_util.py:41-42)
__device__ __forceinline__ float logsubexp_f(float a, float b) {
  float t = -expf(b - a);
  t = (t < -1.0f) ? -1.0f : t;
  return a + log1pf(t);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
  return v;
}

// column sums of a block's old entries accumulated by the contributions
// kernel for the lazy rebalance (all null: not accumulated)
// sums[m] = acc[m] as a float, acc cleared (the end of an order-independent
// column sum; fused.hip)
int colsum_finish(int64_t N, Fx *acc, float *sums, hipStream_t s);

// (order-independent fixed-point sums)
struct OldSums {
  Fx *colsum;            // true values (stored value x column factor)
  Fx *stored;            // stored values
  int32_t *unsafe_reset; // cleared by the launch
};

int posterior_contribs_lazy(const int32_t *rowptr, const int32_t *col, const float *value,
                            const float *colscale, const float *Xn, const float *Xm, int64_t nb,
                            int64_t D, float nu0, float *S, float *nu, Fx *old_fx, Fx *U_old,
                            int32_t *unsafe, hipStream_t s);

// one block's posterior (post_small.hip): [kNN prior covariance] +
// contributions of the block's CSR [+ lazy old column sums] + factorisation
int block_posterior(const lakde_posterior_args &post, const float *X, int64_t N, int64_t row0,
                    const int32_t *rowptr, const int32_t *col, const float *val,
                    const float *colscale, const void *graph, int graph_dtype, int64_t K,
                    float nu0, Fx *old_fx, Fx *U_old, int32_t *unsafe, hipStream_t s,
                    const int32_t *abort = nullptr);

// Packed record layout (floats): [ Linv packed lower row-major: T = D(D+1)/2 ]
// [ x_n: D ] [ estep_const, estep_expo, pred_coeff, pred_expo ], padded to 4.
__host__ __device__ constexpr int64_t rec_tri(int64_t D) { return D * (D + 1) / 2; }
__host__ __device__ constexpr int64_t rec_floats(int64_t D) {
  return ((rec_tri(D) + D + 4) + 3) / 4 * 4;
}
__host__ __device__ constexpr int64_t rec_x(int64_t D) { return rec_tri(D); }
__host__ __device__ constexpr int64_t rec_c(int64_t D) { return rec_tri(D) + D; }

}  // namespace lakde
