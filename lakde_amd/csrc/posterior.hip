// Batched per-point posterior factorisation: one wavefront (64 lanes) per
// kernel point, D <= 63, everything staged in LDS.
//
// Per point n (lakde/local_adaptive_kde.py:83-96, lakde/linalg_util.py:109-170,
// lakde/local_knn_kde.py:195-207, lakde/local_hierarchical_kde.py:269-271):
//   kNN prior: factor prior[n] with its persistent jitter (FP32), write the
//              factor, prior_j = prior[n] + j0 I;  shared prior: prior_j = prior.
//   W_inv = fl(prior_j * fl(nu0 * fl(a/b))) + S[n], factored with the
//   persistent jitter schedule (decisions in FP32, LAPACK potrf semantics:
//   a pivot that is not > 0 fails); a jittered factor is refactored and
//   inverted in FP64, others in FP32.  Then L^{-1}, logdet_W = 2 sum log diag
//   accumulated in FP64, the trace term and the packed record.
#include "record.hpp"

namespace lakde {

constexpr int kPostMaxD = 63;

// leading dimension of the LDS matrices: the smallest odd value >= D
__host__ __device__ constexpr int post_ld(int D) { return D | 1; }

// Reads src[0, count) with kLoadBatch loads in flight per lane and hands
// each value to put(t, v): a copy loop that waited for every load would
// spend a global-memory latency per element.
constexpr int kLoadBatch = 16;

template <typename Put>
__device__ __forceinline__ void load_batched(const float *__restrict__ src, int count, int lane,
                                             Put put) {
  for (int t0 = 0; t0 < count; t0 += kLoadBatch * kWave) {
    float v[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int t = t0 + u * kWave + lane;
      v[u] = t < count ? src[t] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int t = t0 + u * kWave + lane;
      if (t < count) put(t, v[u]);
    }
  }
}

// Lower Cholesky of M (+ jitter on the diagonal, added in FP32) into Lout.
// Lane i owns row i.  Returns true on success (every pivot > 0).
template <typename T>
__device__ bool chol_wave(const float *M, T *Lout, int D, int ld, float jitter) {
  const int lane = threadIdx.x;
  for (int j = 0; j < D; ++j) {
    T s = T(0);
    if (lane >= j && lane < D) {
      float mij = M[lane * ld + j];
      if (lane == j) mij = mij + jitter;
      T acc = T(mij);
      // unrolled so the LDS reads of 8 steps are in flight together (the
      // multiply-adds stay in order)
#pragma unroll 8
      for (int k = 0; k < j; ++k) acc = fma(-Lout[lane * ld + k], Lout[j * ld + k], acc);
      s = acc;
    }
    T piv = __shfl(s, j, kWave);
    if (!(piv > T(0))) {
      __syncthreads();
      return false;
    }
    T d = sqrt(piv);
    if (lane == j) Lout[j * ld + j] = d;
    else if (lane > j && lane < D) Lout[lane * ld + j] = s / d;
    __syncthreads();
  }
  return true;
}

// X = L^{-1} (lower).  Lane c owns column c; rows run in the same order on
// every lane so the L reads are LDS broadcasts.
template <typename T>
__device__ void tril_inv_wave(const T *L, T *X, int D, int ld) {
  const int c = threadIdx.x;
  if (c < D) {
    for (int i = 0; i < D; ++i) {
      T acc = (i == c) ? T(1) : T(0);
#pragma unroll 8
      for (int k = 0; k < i; ++k) acc = fma(-L[i * ld + k], X[k * ld + c], acc);
      X[i * ld + c] = (i < c) ? T(0) : acc / L[i * ld + i];
    }
  }
  __syncthreads();
}

// Factor with the persistent-jitter schedule of stabilized_cholesky:
// try the stored jitter; on failure restart from 0 and add eps * 2^k
// (k = 0..31, accumulated in FP32) until the factorisation succeeds.
// Returns the final jitter; *ok = false when it never converged.
__device__ float stabilized_chol_wave(const float *M, float *L, int D, int ld, float jitter_in,
                                      bool *ok) {
  if (chol_wave<float>(M, L, D, ld, jitter_in)) { *ok = true; return jitter_in; }
  const float eps = 1.1920928955078125e-07f;  // FLT_EPSILON
  float j = 0.f;
  for (int k = 0; k < 32; ++k) {
    j = j + ldexpf(eps, k);
    if (chol_wave<float>(M, L, D, ld, j)) { *ok = true; return j; }
  }
  *ok = false;
  return j;
}

// packed lower triangles of the kNN prior cache: entry n holds prior_j =
// prior + j0 I, then its Cholesky factor, row-major lower (i, k <= i)
__host__ __device__ constexpr int64_t prior_cache_floats(int D) {
  return (int64_t)D * (D + 1);
}

// The kNN prior's factor of one block with its persistent jitter, written to
// the cache (lakde/local_knn_kde.py:195-207, the first call; the prior and
// its jitter do not change during a fit, so every later call of the
// reference returns the same factor)
__global__ void __launch_bounds__(64) prior_cache_kernel(int D, const float *__restrict__ prior,
                                                        float *__restrict__ prior_jitter,
                                                        float *__restrict__ logdet_prior,
                                                        float *__restrict__ cache,
                                                        int32_t *__restrict__ status) {
  extern __shared__ float smem[];
  const int ld = post_ld(D);
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x;
  float *R0 = smem, *R2 = smem + D * ld;
  const int64_t DD = (int64_t)D * D;
  load_batched(prior + n * DD, D * D, lane,
               [&](int t, float v) { R0[(t / D) * ld + (t % D)] = v; });
  __syncthreads();
  bool ok;
  const float j0 = stabilized_chol_wave(R0, R2, D, ld, prior_jitter[n], &ok);
  if (!ok && lane == 0) *status = 1;
  if (lane == 0) prior_jitter[n] = j0;
  double ld0 = lane < D ? log((double)R2[lane * ld + lane]) : 0.0;
  ld0 = wave_sum(ld0);
  if (lane == 0) logdet_prior[n] = (float)(2.0 * ld0);
  const int P = D * (D + 1) / 2;
  float *c = cache + n * prior_cache_floats(D);
  for (int t = lane; t < P; t += kWave) {
    int i, k;
    tri_index(t, &i, &k);
    const float v = R0[i * ld + k];
    c[t] = i == k ? v + j0 : v;
    c[P + t] = R2[i * ld + k];
  }
}

__global__ void __launch_bounds__(64) posterior_kernel(lakde_posterior_args a) {
  extern __shared__ float smem[];
  const int D = (int)a.D;
  // odd row stride: lane i's row-wise reads (L[i * ld + k]) then fall in
  // different LDS banks (an even stride, 64 at D = 63, serialises them)
  const int ld = post_ld(D);
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x;
  float *R0 = smem;               // working matrix
  float *R1 = R0 + D * ld;
  float *R2 = R1 + D * ld;
  float *R3 = R2 + D * ld;
  const int64_t DD = (int64_t)D * D;

  // ---- prior
  const bool cached = a.prior_mode == LAKDE_PRIOR_KNN_CACHED;
  const int Pt = D * (D + 1) / 2;
  if (cached) {
    // prior_j, packed lower (the factorisation reads the lower triangle)
    load_batched(a.prior + n * prior_cache_floats(D), Pt, lane, [&](int t, float v) {
      int i, k;
      tri_index(t, &i, &k);
      R0[i * ld + k] = v;
    });
  } else {
    const float *P = (a.prior_mode == LAKDE_PRIOR_KNN) ? a.prior + n * DD : a.prior;
    load_batched(P, D * D, lane, [&](int t, float v) { R0[(t / D) * ld + (t % D)] = v; });
  }
  __syncthreads();
  if (a.prior_mode == LAKDE_PRIOR_KNN) {
    bool ok;
    float j0 = stabilized_chol_wave(R0, R2, D, ld, a.prior_jitter[n], &ok);
    if (!ok && lane == 0) *a.status = 1;
    if (lane == 0) a.prior_jitter[n] = j0;
    double ld0 = 0.0;
#pragma unroll 4
    for (int t = lane; t < D * D; t += kWave) {
      int i = t / D, k = t % D;
      a.prior_tril[n * DD + t] = (k <= i) ? R2[i * ld + k] : 0.f;
    }
    if (lane < D) ld0 = log((double)R2[lane * ld + lane]);
    ld0 = wave_sum(ld0);
    if (lane == 0) a.logdet_prior[n] = (float)(2.0 * ld0);
    if (lane < D) R0[lane * ld + lane] = R0[lane * ld + lane] + j0;
    __syncthreads();
  }

  // ---- W_inv = prior_j * (nu0 * E[tau]) + S
  const float etau = a.a[n] / a.b[n];
  const float scale = a.nu0 * etau;
  const float *S = a.S + n * DD;
  load_batched(S, D * D, lane, [&](int t, float v) {
    const int i = t / D, k = t % D;
    R0[i * ld + k] = __fadd_rn(__fmul_rn(R0[i * ld + k], scale), v);
  });
  __syncthreads();

  // ---- factor with the persistent jitter (FP32 decisions)
  bool ok;
  float jw = stabilized_chol_wave(R0, R2, D, ld, a.W_jitter[n], &ok);
  if (!ok && lane == 0) *a.status = 1;
  if (lane == 0) a.W_jitter[n] = jw;

  // Linv ends in R1 (FP32 path) or R2 (FP64 path); R3 is free afterwards.
  float *Linv;
  if (jw != 0.f && ok) {
    // near-singular: refactor fl32(W_inv + j I) and invert in FP64
    double *L64 = reinterpret_cast<double *>(R2);  // overlays R2|R3
    double *I64 = reinterpret_cast<double *>(R0);  // overlays R0|R1
    if (chol_wave<double>(R0, L64, D, ld, jw)) {
      tril_inv_wave<double>(L64, I64, D, ld);
      for (int t = lane; t < D * ld; t += kWave) R2[t] = (float)I64[t];
      __syncthreads();
      Linv = R2;
    } else {
      // FP64 disagrees with the FP32 decision: keep the FP32 factor
      chol_wave<float>(R0, R2, D, ld, jw);
      tril_inv_wave<float>(R2, R1, D, ld);
      Linv = R1;
    }
  } else {
    tril_inv_wave<float>(R2, R1, D, ld);
    Linv = R1;
  }

  // ---- outputs: Linv, logdet_W
#pragma unroll 4
  for (int t = lane; t < D * D; t += kWave) {
    int i = t / D, k = t % D;
    a.Linv[n * DD + t] = (k <= i) ? Linv[i * ld + k] : 0.f;
  }
  double lsum = (lane < D) ? log((double)Linv[lane * ld + lane]) : 0.0;
  lsum = wave_sum(lsum);
  const float logdet = (float)(2.0 * lsum);

  // ---- trace = nu * sum_{i,j} (sum_k T[k][i] Linv[j][k])^2
  // Linv is lower triangular (k <= j).  The kNN prior's T is its Cholesky
  // factor, lower triangular too, so only i <= k <= j contribute there: the
  // terms left out are exact zeros (the sums are unchanged for finite
  // factors) and the longest lane runs D(D+1)/2 steps instead of D^2.  The
  // shared prior's T is the dense prior covariance (all i, all k <= j).
  float *Tm = R3;
  const bool tri = a.prior_mode != LAKDE_PRIOR_SHARED;
  __syncthreads();
  if (cached) {
    load_batched(a.prior + n * prior_cache_floats(D) + Pt, Pt, lane, [&](int t, float v) {
      int i, k;
      tri_index(t, &i, &k);
      Tm[i * ld + k] = v;
    });
  } else {
    const float *Tsrc = tri ? a.prior_tril + n * DD : a.prior;
    load_batched(Tsrc, D * D, lane, [&](int t, float v) { Tm[(t / D) * ld + (t % D)] = v; });
  }
  __syncthreads();
  float tr = 0.f;
  if (lane < D) {
    const int iend = tri ? lane + 1 : D;
    for (int i = 0; i < iend; ++i) {
      float s = 0.f;
#pragma unroll 8
      for (int k = tri ? i : 0; k <= lane; ++k) s = fmaf(Tm[k * ld + i], Linv[lane * ld + k], s);
      tr = fmaf(s, s, tr);
    }
  }
  tr = wave_sum(tr);
  const float nu = a.nu[n];

  // ---- packed record
  const int64_t R = rec_floats(D);
  float *rec = a.rec + n * R;
#pragma unroll 4
  for (int t = lane; t < (int)rec_tri(D); t += kWave) {
    int i, r;
    tri_index(t, &i, &r);
    rec[t] = Linv[i * ld + r];
  }
  if (lane < D) rec[rec_x(D) + lane] = a.Xn[n * D + lane];
  if (lane == 0) {
    a.logdet_W[n] = logdet;
    const float trace = nu * tr;
    a.trace[n] = trace;
    if (a.b_out != nullptr) a.b_out[n] = a.b0 + (a.nu0 * trace) / 2.f;
    record_consts(rec + rec_c(D), nu, logdet, D);
  }
}

}  // namespace lakde

using namespace lakde;

extern "C" int lakde_posterior(const lakde_posterior_args *args, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr, "null args");
  const lakde_posterior_args &a = *args;
  LAKDE_CHECK_ARG(a.D >= 1 && a.D <= kPostMaxD, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(a.nb >= 0, "negative block size");
  LAKDE_CHECK_ARG(a.prior && a.S && a.nu && a.a && a.b && a.W_jitter && a.Xn && a.Linv &&
                      a.logdet_W && a.trace && a.rec && a.status,
                  "null pointer");
  LAKDE_CHECK_ARG(a.prior_mode == LAKDE_PRIOR_KNN || a.prior_mode == LAKDE_PRIOR_SHARED ||
                      a.prior_mode == LAKDE_PRIOR_KNN_CACHED,
                  "bad prior mode");
  if (a.prior_mode == LAKDE_PRIOR_KNN)
    LAKDE_CHECK_ARG(a.prior_jitter && a.prior_tril && a.logdet_prior, "kNN prior outputs missing");
  if (a.nb == 0) return LAKDE_OK;
  size_t lds = (size_t)4 * a.D * post_ld((int)a.D) * sizeof(float);
  hipLaunchKernelGGL(posterior_kernel, dim3((unsigned)a.nb), dim3(64), lds, as_stream(stream), a);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

extern "C" int64_t lakde_prior_cache_floats(int64_t D) {
  return (D >= 1 && D <= kPostMaxD) ? prior_cache_floats((int)D) : -1;
}

extern "C" int lakde_knn_prior_cache(const float *X, int64_t N, int64_t D, int64_t row0,
                                     int64_t nb, const void *graph, int graph_dtype, int64_t K,
                                     float *prior_scratch, float *prior_jitter,
                                     float *logdet_prior, float *cache, int32_t *status,
                                     lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 1 && D <= kPostMaxD, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(X && graph && prior_scratch && prior_jitter && logdet_prior && cache && status,
                  "null pointer");
  if (nb == 0) return LAKDE_OK;
  int rc = lakde_knn_prior_cov(X, N, D, row0, nb, graph, graph_dtype, K, prior_scratch, stream);
  if (rc) return rc;
  size_t lds = (size_t)2 * D * post_ld((int)D) * sizeof(float);
  hipLaunchKernelGGL(prior_cache_kernel, dim3((unsigned)nb), dim3(64), lds, as_stream(stream),
                     (int)D, prior_scratch, prior_jitter, logdet_prior, cache, status);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}
