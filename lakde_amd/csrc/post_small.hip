// Fused posterior of one kernel block for small dimensions (D <= 6): the
// kNN prior covariance, the contributions S_n and nu_n of the block's
// responsibilities (with the lazy rebalance's old column sums), the two
// jittered factorisations, L^{-1}, log-determinants, the trace term, the tau
// update and the packed record, from one launch.  One workgroup of four
// wavefronts per kernel point: lanes gather the neighbours / entries,
// butterfly reductions leave the D(D+1)/2 sums in every lane, the waves'
// partial sums are combined through LDS, and every lane of wave 0 then runs
// the same register-resident factorisation, so nothing intermediate goes
// through HBM.  For larger D the step uses lakde_knn_prior_cov,
// lakde_posterior_contribs and lakde_posterior (posterior.hip), whose
// arithmetic this kernel follows operation by operation: the same FP32
// jitter decisions, the same FP64 refactorisation of jittered factors.
//
// Reference: lakde/utils.py:40-65 (batched_knn_cov), lakde/kernels.py:26-39
// and lakde/local_adaptive_kde.py:66-96 (contributions, posterior),
// lakde/linalg_util.py:30-37,49-52,109-170 (logdet, trace, tril_inverse,
// stabilized_cholesky), lakde/local_knn_kde.py:195-207,220-236.
#include "record.hpp"

namespace lakde {

constexpr int kSmallMaxD = 6;

struct SmallPost {
  lakde_posterior_args post;
  const float *X;
  int64_t row0;
  const int32_t *rowptr, *col;  // the block's CSR (row pointer local to the block)
  const float *val, *colscale;  // colscale: per-column factors of the stored values, or null
  const void *graph;            // kNN rows of the block (nb x K), or null for a shared prior
  int graph_dtype;
  int64_t K;
  float nu0;                    // nu = nu0 + sum of the row's responsibilities
  Fx *old_colsum;               // optional lazy rebalance old column sums (true values)
  Fx *U_old;                    //   and stored values (order-independent sums)
  int32_t *unsafe;              //   unsafe-step flag, cleared by the launch
  const int32_t *abort;         // optional: the launch does nothing when *abort != 0
};

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

template <int P>
__device__ __forceinline__ void wave_sum_all(float (&a)[P]) {
  // IEEE addition commutes, so both partners of every exchange form the
  // same sum and all lanes end with identical bits
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int t = 0; t < P; ++t) a[t] += __shfl_xor(a[t], off, kWave);
  }
}

// Lower Cholesky of M + jitter I (M: packed lower triangle of a symmetric
// matrix), in the operation order of chol_wave (posterior.hip): the jitter
// is added to the diagonal in FP32, then the column is formed in T.  Returns
// false at the first pivot that is not > 0 (LAPACK potrf semantics).
template <typename T, int D>
__device__ __forceinline__ bool chol_reg(const float (&M)[D * (D + 1) / 2], float jitter,
                                         T (&L)[D * (D + 1) / 2]) {
#pragma unroll
  for (int j = 0; j < D; ++j) {
    T piv = T(M[tri(j, j)] + jitter);
#pragma unroll
    for (int k = 0; k < j; ++k) piv = fma(-L[tri(j, k)], L[tri(j, k)], piv);
    if (!(piv > T(0))) return false;
    const T d = sqrt(piv);
    L[tri(j, j)] = d;
#pragma unroll
    for (int i = j + 1; i < D; ++i) {
      T acc = T(M[tri(i, j)]);
#pragma unroll
      for (int k = 0; k < j; ++k) acc = fma(-L[tri(i, k)], L[tri(j, k)], acc);
      L[tri(i, j)] = acc / d;
    }
  }
  return true;
}

// X = L^{-1}, lower, column by column as tril_inv_wave (the terms that
// tril_inv_wave multiplies by the zeros above the diagonal are skipped;
// they leave its sums unchanged)
template <typename T, int D>
__device__ __forceinline__ void tril_inv_reg(const T (&L)[D * (D + 1) / 2],
                                             T (&X)[D * (D + 1) / 2]) {
#pragma unroll
  for (int c = 0; c < D; ++c) {
    X[tri(c, c)] = T(1) / L[tri(c, c)];
#pragma unroll
    for (int i = c + 1; i < D; ++i) {
      T acc = T(0);
#pragma unroll
      for (int k = c; k < i; ++k) acc = fma(-L[tri(i, k)], X[tri(k, c)], acc);
      X[tri(i, c)] = acc / L[tri(i, i)];
    }
  }
}

// sum_i log(L_ii) in FP64, the logs computed one per lane and summed in
// index order (every lane holds the same factor and gets the same sum)
template <int D>
__device__ __forceinline__ double diag_log_sum(const float (&L)[D * (D + 1) / 2], int lane) {
  float v = 1.f;
#pragma unroll
  for (int i = 0; i < D; ++i) v = lane == i ? L[tri(i, i)] : v;
  const double lg = log((double)v);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) s += __shfl(lg, i, kWave);
  return s;
}

// stabilized_cholesky's schedule: the stored jitter first; on failure
// restart from 0 and add eps * 2^k (k = 0..31, accumulated in FP32)
template <int D>
__device__ float stab_chol_reg(const float (&M)[D * (D + 1) / 2], float jitter_in,
                               float (&L)[D * (D + 1) / 2], bool *ok) {
  if (chol_reg<float, D>(M, jitter_in, L)) { *ok = true; return jitter_in; }
  const float eps = 1.1920928955078125e-07f;  // FLT_EPSILON
  float j = 0.f;
  for (int k = 0; k < 32; ++k) {
    j = j + ldexpf(eps, k);
    if (chol_reg<float, D>(M, j, L)) { *ok = true; return j; }
  }
  *ok = false;
  return j;
}

// waves per kernel point: all share the row's entries, wave 0 also builds
// the kNN prior and then runs the factorisation
constexpr int kPostWaves = 4;

template <int D>
__global__ void __launch_bounds__(kPostWaves * kWave) post_small_kernel(SmallPost p) {
  constexpr int P = D * (D + 1) / 2;
  __shared__ float sm_part[kPostWaves][P + 1];
  const lakde_posterior_args &a = p.post;
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  // a posterior enqueued ahead of the previous step's flags: that step needs
  // the host, which enqueues this one again after handling it
  if (p.abort != nullptr && *p.abort != 0) return;
  if (p.unsafe != nullptr && n == 0 && threadIdx.x == 0) *p.unsafe = 0;
  float xn[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xn[j] = p.X[(p.row0 + n) * D + j];

  // ---- prior covariance (lakde/utils.py:40-65, self-centred, / K), wave 0
  float pri[P];
  const bool knn = a.prior_mode == LAKDE_PRIOR_KNN;
  if (knn && w == 0) {
#pragma unroll
    for (int t = 0; t < P; ++t) pri[t] = 0.f;
    for (int64_t k = lane; k < p.K; k += kWave) {
      const int64_t g = p.graph_dtype == LAKDE_I32 ? (int64_t)((const int32_t *)p.graph)[n * p.K + k]
                                                   : ((const int64_t *)p.graph)[n * p.K + k];
      float d[D];
#pragma unroll
      for (int j = 0; j < D; ++j) d[j] = p.X[g * D + j] - xn[j];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) pri[tri(i, j)] = fmaf(d[i], d[j], pri[tri(i, j)]);
    }
    wave_sum_all<P>(pri);
    const float fk = (float)p.K;
#pragma unroll
    for (int t = 0; t < P; ++t) pri[t] = pri[t] / fk;
  } else if (!knn) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) pri[tri(i, j)] = a.prior[i * D + j];
  }

  // ---- contributions S_n = sum_e r_e d_e d_e^T, nu_n = nu0 + sum_e r_e
  // (lakde/kernels.py:26-39, lakde/local_adaptive_kde.py:66-81).  The
  // waves take EB wavefronts of entries each per pass, and all loads of a
  // pass are issued before any is used: the long rows (kernels with many
  // kept points) set the launch's time.  Partial sums are combined in a
  // fixed order (deterministic, not the reference's sequential order).
  float S[P];
#pragma unroll
  for (int t = 0; t < P; ++t) S[t] = 0.f;
  float wsum = 0.f;
  const int32_t beg = p.rowptr[n], end = p.rowptr[n + 1];
  constexpr int EB = 4;
  for (int32_t e0 = beg + w * EB * kWave; e0 < end; e0 += kPostWaves * EB * kWave) {
    int32_t c[EB];
    float u[EB], g[EB], d[EB][D];
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int32_t e = e0 + i * kWave + lane;
      c[i] = e < end ? p.col[e] : -1;
      u[i] = e < end ? p.val[e] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      g[i] = (c[i] >= 0 && p.colscale != nullptr) ? p.colscale[c[i]] : 1.f;
#pragma unroll
      for (int j = 0; j < D; ++j) d[i][j] = c[i] >= 0 ? p.X[(int64_t)c[i] * D + j] - xn[j] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const float wt = p.colscale != nullptr ? u[i] * g[i] : u[i];
      if (c[i] >= 0 && p.old_colsum != nullptr) {
        fx_atomic_add(&p.old_colsum[c[i]], wt);
        fx_atomic_add(&p.U_old[c[i]], u[i]);
      }
#pragma unroll
      for (int ii = 0; ii < D; ++ii) {
        const float wd = wt * d[i][ii];
#pragma unroll
        for (int j = 0; j <= ii; ++j) S[tri(ii, j)] = fmaf(wd, d[i][j], S[tri(ii, j)]);
      }
      wsum += wt;
    }
  }
  wave_sum_all<P>(S);
  wsum = wave_sum(wsum);
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < P; ++t) sm_part[w][t] = S[t];
    sm_part[w][P] = wsum;
  }
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int t = 0; t < P; ++t) S[t] = ((sm_part[0][t] + sm_part[1][t]) + sm_part[2][t]) + sm_part[3][t];
  const float nu =
      p.nu0 + (((sm_part[0][P] + sm_part[1][P]) + sm_part[2][P]) + sm_part[3][P]);

  // ---- prior factor with its persistent jitter (kNN)
  float T[P];  // lower factor of the prior (kNN); the shared prior is read in the trace
  if (knn) {
    bool ok;
    const float j0 = stab_chol_reg<D>(pri, a.prior_jitter[n], T, &ok);
    const double ld0 = diag_log_sum<D>(T, lane);
    if (lane == 0) {
      if (!ok) *a.status = 1;
      a.prior_jitter[n] = j0;
      a.logdet_prior[n] = (float)(2.0 * ld0);
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) a.prior_tril[(n * D + i) * D + j] = j <= i ? T[tri(i, j)] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) pri[tri(i, i)] = pri[tri(i, i)] + j0;
  }

  // ---- W_inv = prior_j * (nu0 E[tau]) + S, factored with its jitter
  const float etau = a.a[n] / a.b[n];
  const float scale = a.nu0 * etau;
  float W[P];
#pragma unroll
  for (int t = 0; t < P; ++t) W[t] = __fadd_rn(__fmul_rn(pri[t], scale), S[t]);
  bool ok;
  float L[P];
  const float jw = stab_chol_reg<D>(W, a.W_jitter[n], L, &ok);
  float Li[P];
  if (jw != 0.f && ok) {
    // near-singular: refactor fl32(W_inv + j I) and invert in FP64
    double L64[P], I64[P];
    if (chol_reg<double, D>(W, jw, L64)) {
      tril_inv_reg<double, D>(L64, I64);
#pragma unroll
      for (int t = 0; t < P; ++t) Li[t] = (float)I64[t];
    } else {
      chol_reg<float, D>(W, jw, L);
      tril_inv_reg<float, D>(L, Li);
    }
  } else {
    tril_inv_reg<float, D>(L, Li);
  }

  // ---- logdet, trace = nu ||T^T Linv^T||_F^2 (lakde/linalg_util.py:30-37)
  const float logdet = (float)(2.0 * diag_log_sum<D>(Li, lane));
  float tr = 0.f;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    float trc = 0.f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float s = 0.f;
      if (knn) {
#pragma unroll
        for (int k = i; k <= c; ++k) s = fmaf(T[tri(k, i)], Li[tri(c, k)], s);
      } else {
#pragma unroll
        for (int k = 0; k <= c; ++k) s = fmaf(a.prior[k * D + i], Li[tri(c, k)], s);
      }
      trc = fmaf(s, s, trc);
    }
    tr += trc;
  }

  if (lane == 0) {
    if (!ok) *a.status = 1;
    a.W_jitter[n] = jw;
    const_cast<float *>(a.nu)[n] = nu;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) a.Linv[(n * D + i) * D + j] = j <= i ? Li[tri(i, j)] : 0.f;
    a.logdet_W[n] = logdet;
    const float trace = nu * tr;
    a.trace[n] = trace;
    if (a.b_out != nullptr) a.b_out[n] = a.b0 + (a.nu0 * trace) / 2.f;
    float *rec = a.rec + n * rec_floats(D);
#pragma unroll
    for (int t = 0; t < P; ++t) rec[t] = Li[t];
#pragma unroll
    for (int j = 0; j < D; ++j) rec[rec_x(D) + j] = xn[j];
  }
  record_consts_wave(a.rec + n * rec_floats(D) + rec_c(D), nu, logdet, D, lane);
}

template <int D>
static int launch_post_small(const SmallPost &p, hipStream_t s) {
  hipLaunchKernelGGL((post_small_kernel<D>), dim3((unsigned)p.post.nb), dim3(kPostWaves * kWave),
                     0, s, p);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// One block's posterior: the fused kernel for D <= kSmallMaxD, otherwise the
// kNN prior, contributions and factorisation kernels in turn.
int block_posterior(const lakde_posterior_args &post, const float *X, int64_t N, int64_t row0,
                    const int32_t *rowptr, const int32_t *col, const float *val,
                    const float *colscale, const void *graph, int graph_dtype, int64_t K,
                    float nu0, Fx *old_colsum, Fx *U_old, int32_t *unsafe, hipStream_t s,
                    const int32_t *abort) {
  const int64_t D = post.D, nb = post.nb;
  LAKDE_CHECK_ARG(D >= 1 && D <= 63 && nb >= 0 && row0 >= 0 && row0 + nb <= N, "bad sizes");
  LAKDE_CHECK_ARG(X && rowptr && col && val, "null pointer");
  LAKDE_CHECK_ARG((post.prior_mode == LAKDE_PRIOR_KNN) == (graph != nullptr),
                  "a kNN prior needs the graph rows, a shared prior none");
  LAKDE_CHECK_ARG(graph == nullptr || graph_dtype == LAKDE_I32 || graph_dtype == LAKDE_I64,
                  "graph must be int32 or int64");
  LAKDE_CHECK_ARG((old_colsum == nullptr) == (U_old == nullptr), "old sums come in pairs");
  LAKDE_CHECK_ARG(post.prior_mode != LAKDE_PRIOR_KNN_CACHED || D > kSmallMaxD,
                  "the cached kNN prior needs D >= 7");
  if (nb == 0) return LAKDE_OK;
  if (D <= kSmallMaxD) {
    LAKDE_CHECK_ARG(post.prior_mode != LAKDE_PRIOR_KNN ||
                        (post.prior_jitter && post.prior_tril && post.logdet_prior),
                    "kNN prior outputs missing");
    LAKDE_CHECK_ARG(post.prior && post.nu && post.a && post.b && post.W_jitter && post.Linv &&
                        post.logdet_W && post.trace && post.rec && post.status,
                    "null pointer");
    SmallPost p{post, X, row0, rowptr, col, val, colscale, graph, graph_dtype, K, nu0,
                old_colsum, U_old, unsafe, abort};
    switch (D) {
      case 1: return launch_post_small<1>(p, s);
      case 2: return launch_post_small<2>(p, s);
      case 3: return launch_post_small<3>(p, s);
      case 4: return launch_post_small<4>(p, s);
      case 5: return launch_post_small<5>(p, s);
      case 6: return launch_post_small<6>(p, s);
    }
  }
  LAKDE_CHECK_ARG(abort == nullptr, "a guarded posterior needs D <= 6");
  int rc;
  if (graph != nullptr) {
    rc = lakde_knn_prior_cov(X, N, D, row0, nb, graph, graph_dtype, K,
                             const_cast<float *>(post.prior), s);
    if (rc) return rc;
  }
  if (old_colsum != nullptr) {
    rc = posterior_contribs_lazy(rowptr, col, val, colscale, X + row0 * D, X, nb, D, nu0,
                                 const_cast<float *>(post.S), const_cast<float *>(post.nu),
                                 old_colsum, U_old, unsafe, s);
  } else {
    rc = lakde_posterior_contribs(rowptr, LAKDE_I32, col, LAKDE_I32, val, colscale, X + row0 * D,
                                  X, nb, N, D, nu0, const_cast<float *>(post.S),
                                  const_cast<float *>(post.nu), s);
  }
  if (rc) return rc;
  return lakde_posterior(&post, s);
}

}  // namespace lakde

using namespace lakde;

extern "C" int lakde_block_posterior(const lakde_posterior_args *post, const float *X, int64_t N,
                                     int64_t row0, const int32_t *rowptr, const int32_t *col,
                                     const float *val, const float *colscale,
                                     const void *knn_graph, int knn_graph_dtype, int64_t knn_k,
                                     lakde_stream_t stream) {
  LAKDE_CHECK_ARG(post != nullptr, "null args");
  return block_posterior(*post, X, N, row0, rowptr, col, val, colscale, knn_graph,
                         knn_graph_dtype, knn_k, post->nu0, nullptr, nullptr, nullptr,
                         as_stream(stream));
}
