// The three reference ops (ind2ptr, W_inv, chol_quad_form) plus the block
// helpers built on the same segmented outer-product kernel (kNN prior
// covariance, posterior contributions), written for gfx950 wave64.
#include <mutex>

#include <hip/hip_fp16.h>

#include "common.hpp"

namespace lakde {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

// ------------------------------------------------------------------ ind2ptr
// rowptr[i] = lower_bound(ind, i): one thread per output entry, binary search
// over the sorted index array.  Same result as the run-length/cumsum
// formulation of lakde/kernels.py:11-17 for sorted input.
template <typename In, typename Out>
__global__ void __launch_bounds__(256) ind2ptr_kernel(const In *__restrict__ ind, int64_t nnz,
                                                     int64_t M, Out *__restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i > M) return;
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)ind[mid] < i) lo = mid + 1; else hi = mid;
  }
  out[i] = (Out)lo;
}

template <typename In, typename Out>
static int launch_ind2ptr(const void *ind, int64_t nnz, int64_t M, void *out, hipStream_t s) {
  int64_t n = M + 1;
  hipLaunchKernelGGL((ind2ptr_kernel<In, Out>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                     (const In *)ind, nnz, M, (Out *)out);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// ------------------------------------------------- segmented outer products
// One workgroup (256 threads) per output row n.  The row's entries are staged
// through LDS in chunks of E rows of d_e = Xm[col_e] - Xn[n] (coalesced row
// gathers); each thread owns a 3 x 3 tile of the upper triangle (23 <= D <=
// 63) or up to kMaxPairs (i <= j) output pairs and, when the triangle is
// small, thread groups split the entries and are reduced in a fixed order.
// No atomics; the result is deterministic.
//   GRAPH = false: CSR rows (rowptr, col, value weights)       -> W_inv
//   GRAPH = true : dense K-regular graph, unit weights, / K    -> kNN prior
constexpr int kOuterThreads = 256;
constexpr int kMaxPairs = 9;          // 256 * 9 >= 63*64/2 + 64
constexpr int kStageElems = 4096;     // LDS staging elements per chunk
constexpr int kStageBatch = 8;        // staging loads in flight per thread
// 3 x 3 register tiles of the upper triangle for 23 <= D <= 63: one tile per
// thread (ceil(D/3) (ceil(D/3) + 1) / 2 <= 231 tiles), 6 LDS reads per
// entry for 9 products instead of 2 per product.  Every output pair i <= j
// is still fma(w * d_i, d_j, acc) over the entries in order, so the sums are
// the same as with the pair-per-slot mapping.
constexpr int kTileMinD = 23;         // below: P <= 256, the entry-group split
constexpr int kTileMaxD = 63;

// S: storage type of value, Xn, Xm and out (the arithmetic is in T; S =
// __half with T = float for the float16 form of W_inv)
template <typename T, typename RP, typename CI, bool GRAPH, typename S = T>
__global__ void __launch_bounds__(kOuterThreads)
outer_sum_kernel(const RP *__restrict__ rowptr, const CI *__restrict__ col,
                 const S *__restrict__ value, const S *__restrict__ Xn,
                 const S *__restrict__ Xm, int64_t D, int64_t K, S *__restrict__ out,
                 T *__restrict__ nu_out, T nu0, const T *__restrict__ colscale,
                 OldSums old) {
  __shared__ T stage[kStageElems];
  __shared__ T wts[64];
  __shared__ T xn_s[512];
  const int64_t n = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t P = D * (D + 1) / 2;
  const int64_t E = (kStageElems / D) < 64 ? (kStageElems / D) : 64;

  int64_t beg, end;
  if (GRAPH) { beg = n * K; end = beg + K; }
  else { beg = (int64_t)rowptr[n]; end = (int64_t)rowptr[n + 1]; }

  for (int64_t d = tid; d < D; d += kOuterThreads) xn_s[d] = (T)Xn[n * D + d];
  if (old.unsafe_reset != nullptr && n == 0 && tid == 0) *old.unsafe_reset = 0;
  T nu_acc = nu0;

  // tile of this thread (tiled form): rows 3 ti .., columns 3 tj .., ti <= tj
  const bool tiled = D >= kTileMinD && D <= kTileMaxD;
  const int nT = (int)(D + 2) / 3;
  int ti = 0, tj = 0;
  const bool tile_ok = tiled && tid < nT * (nT + 1) / 2;
  if (tile_ok) {
    int r = tid;
    while (r >= nT - ti) { r -= nT - ti; ++ti; }
    tj = ti + r;
  }
  // tile rows/columns clamped into [0, D) for the reads (entries outside the
  // triangle are computed and not written)
  int ir[3], jc[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    ir[u] = 3 * ti + u < D ? 3 * ti + u : (int)D - 1;
    jc[u] = 3 * tj + u < D ? 3 * tj + u : (int)D - 1;
  }

  // pair batches (a single batch for D <= 70)
  for (int64_t pbase = 0; pbase < P; pbase += (int64_t)kOuterThreads * kMaxPairs) {
    const int64_t Pb = (P - pbase) < (int64_t)kOuterThreads * kMaxPairs
                           ? (P - pbase) : (int64_t)kOuterThreads * kMaxPairs;
    // group split of the entries when the triangle is small
    const int G = Pb <= kOuterThreads ? (int)(kOuterThreads / Pb) : 1;
    const int g = Pb <= kOuterThreads ? (int)(tid / Pb) : 0;
    const bool active = g < G;
    int pi[kMaxPairs], pj[kMaxPairs];
    int np = 0;
    T acc[kMaxPairs];
#pragma unroll
    for (int q = 0; q < kMaxPairs; ++q) { acc[q] = T(0); pi[q] = 0; pj[q] = 0; }
    if (active && !tiled) {
      int64_t p0 = Pb <= kOuterThreads ? (tid % Pb) : tid;
      for (int q = 0; q < kMaxPairs; ++q) {
        int64_t p = p0 + (int64_t)q * kOuterThreads;
        if (Pb <= kOuterThreads && q > 0) break;
        if (p >= Pb) break;
        int64_t pp = pbase + p, i = 0;
        while (pp >= D - i) { pp -= D - i; ++i; }
        pi[q] = (int)i;
        pj[q] = (int)(i + pp);
        ++np;
      }
    }
    for (int64_t c0 = beg; c0 < end; c0 += E) {
      const int64_t ne = (end - c0) < E ? (end - c0) : E;
      __syncthreads();
      const int Di = (int)D;
      const int nt = (int)ne * Di;
      // kStageBatch column indices, then the kStageBatch gathered
      // coordinates, in flight per thread (two dependent loads per element)
      for (int t0 = tid; t0 < nt; t0 += kOuterThreads * kStageBatch) {
        int64_t cidx[kStageBatch];
        T xv[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int t = t0 + u * kOuterThreads;
          cidx[u] = t < nt ? (int64_t)col[c0 + t / Di] : 0;
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int t = t0 + u * kOuterThreads;
          xv[u] = t < nt ? (T)Xm[cidx[u] * D + (t - (t / Di) * Di)] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int t = t0 + u * kOuterThreads;
          if (t < nt) stage[t] = xv[u] - xn_s[t - (t / Di) * Di];
        }
      }
      for (int64_t e = tid; e < ne; e += kOuterThreads) {
        T wv = GRAPH ? T(1) : (T)value[c0 + e];
        if (!GRAPH && colscale != nullptr) wv = wv * colscale[(int64_t)col[c0 + e]];
        wts[e] = wv;
        if (!GRAPH && old.colsum != nullptr && pbase == 0) {
          const int64_t ci = (int64_t)col[c0 + e];
          fx_atomic_add(&old.colsum[ci], (float)wv);
          fx_atomic_add(&old.stored[ci], (float)(T)value[c0 + e]);
        }
      }
      __syncthreads();
      // nu = nu0 + sum of the row's weights in entry order, from the staged
      // weights (first pair batch only)
      if (!GRAPH && nu_out != nullptr && pbase == 0 && tid == 0)
        for (int64_t e = 0; e < ne; ++e) nu_acc += wts[e];
      if (tiled) {
        if (tile_ok) {
          for (int64_t e = 0; e < ne; ++e) {
            const T w = wts[e];
            const T *de = stage + e * D;
            T wi[3], dj[3];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
              wi[u] = w * de[ir[u]];
              dj[u] = de[jc[u]];
            }
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
              for (int b2 = 0; b2 < 3; ++b2)
                acc[a2 * 3 + b2] = fma(wi[a2], dj[b2], acc[a2 * 3 + b2]);
          }
        }
      } else if (active) {
        for (int64_t e = g; e < ne; e += G) {
          const T w = wts[e];
          const T *de = stage + e * D;
#pragma unroll
          for (int q = 0; q < kMaxPairs; ++q)
            if (q < np) acc[q] = fma(w * de[pi[q]], de[pj[q]], acc[q]);
        }
      }
    }
    // reduce over groups in a fixed order, write both triangles
    __syncthreads();
    if (G > 1) {
      if (active) stage[tid] = acc[0];
      __syncthreads();
      if (g == 0 && active) {
        T s = stage[tid];
        for (int gg = 1; gg < G; ++gg) s += stage[gg * Pb + tid];
        acc[0] = s;
      }
    }
    if (tile_ok) {
#pragma unroll
      for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 3; ++b2) {
          const int64_t i = 3 * ti + a2, j = 3 * tj + b2;
          if (i < D && j < D && i <= j) {
            T v = acc[a2 * 3 + b2];
            if (GRAPH) v = v / T(K);
            out[(n * D + i) * D + j] = (S)v;
            out[(n * D + j) * D + i] = (S)v;
          }
        }
    } else if (g == 0 && active && !tiled) {
      for (int q = 0; q < np; ++q) {
        T v = acc[q];
        if (GRAPH) v = v / T(K);
        out[(n * D + pi[q]) * D + pj[q]] = (S)v;
        out[(n * D + pj[q]) * D + pi[q]] = (S)v;
      }
    }
  }
  if (!GRAPH && nu_out != nullptr && tid == 0) nu_out[n] = nu_acc;
}

template <typename T, typename RP, typename CI, bool GRAPH, typename S = T>
static int launch_outer(const void *rowptr, const void *col, const void *value, const void *Xn,
                        const void *Xm, int64_t N, int64_t D, int64_t K, void *out, void *nu,
                        T nu0, hipStream_t s, const T *colscale = nullptr,
                        OldSums old = OldSums{nullptr, nullptr, nullptr}) {
  if (N == 0) return LAKDE_OK;
  hipLaunchKernelGGL((outer_sum_kernel<T, RP, CI, GRAPH, S>), dim3((unsigned)N),
                     dim3(kOuterThreads), 0, s, (const RP *)rowptr, (const CI *)col,
                     (const S *)value, (const S *)Xn, (const S *)Xm, D, K, (S *)out, (T *)nu, nu0,
                     colscale, old);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

template <typename T, typename S = T>
static int dispatch_outer_idx(int rp_dt, int col_dt, const void *rowptr, const void *col,
                              const void *value, const void *Xn, const void *Xm, int64_t N,
                              int64_t D, void *out, void *nu, T nu0, hipStream_t s,
                              const T *cs = nullptr) {
  if (rp_dt == LAKDE_I32 && col_dt == LAKDE_I32)
    return launch_outer<T, int32_t, int32_t, false, S>(rowptr, col, value, Xn, Xm, N, D, 0, out, nu, nu0, s, cs);
  if (rp_dt == LAKDE_I32 && col_dt == LAKDE_I64)
    return launch_outer<T, int32_t, int64_t, false, S>(rowptr, col, value, Xn, Xm, N, D, 0, out, nu, nu0, s, cs);
  if (rp_dt == LAKDE_I64 && col_dt == LAKDE_I32)
    return launch_outer<T, int64_t, int32_t, false, S>(rowptr, col, value, Xn, Xm, N, D, 0, out, nu, nu0, s, cs);
  if (rp_dt == LAKDE_I64 && col_dt == LAKDE_I64)
    return launch_outer<T, int64_t, int64_t, false, S>(rowptr, col, value, Xn, Xm, N, D, 0, out, nu, nu0, s, cs);
  set_error("unsupported index dtype combination (int32/int64 only)");
  return LAKDE_EDTYPE;
}

// ------------------------------------------------------------ quad form
// Workgroup = 256 points x NK kernels.  The NK factors (transposed to the
// lower L = W^T, zero padded to DP x DP) and centres are staged in LDS; every
// thread keeps its point's difference vector in registers and streams the
// triangular mat-vec from LDS (broadcast reads, row-contiguous).
template <typename T, int DP>
struct QfCfg {
  static constexpr int raw = 32768 / (DP * DP * (int)sizeof(T));
  static constexpr int NK = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
};

template <typename T, int DP, typename S = T>
__global__ void __launch_bounds__(256)
quad_form_kernel(const S *__restrict__ W, int64_t sW0, int64_t sW1, int64_t sW2,
                 const S *__restrict__ Xn, int64_t sXn0, int64_t sXn1,
                 const S *__restrict__ Xm, int64_t sXm0, int64_t sXm1,
                 S *__restrict__ out, int64_t so0, int64_t so1, int64_t N, int64_t M, int D) {
  constexpr int NK = QfCfg<T, DP>::NK;
  __shared__ T Ls[NK][DP][DP];
  __shared__ T xs[NK][DP];
  const int tid = threadIdx.x;
  const int64_t n0 = (int64_t)blockIdx.y * NK;
  const int64_t m = (int64_t)blockIdx.x * 256 + tid;
  for (int t = tid; t < NK * DP * DP; t += 256) {
    int k = t / (DP * DP), r = t - k * DP * DP, i = r / DP, j = r - i * DP;
    int64_t n = n0 + k;
    T v = T(0);
    if (n < N && i < D && j <= i) v = (T)W[n * sW0 + (int64_t)j * sW1 + (int64_t)i * sW2];
    Ls[k][i][j] = v;
  }
  for (int t = tid; t < NK * DP; t += 256) {
    int k = t / DP, j = t - k * DP;
    int64_t n = n0 + k;
    xs[k][j] = (n < N && j < D) ? (T)Xn[n * sXn0 + (int64_t)j * sXn1] : T(0);
  }
  __syncthreads();
  if (m >= M) return;
  T xm[DP];
#pragma unroll
  for (int j = 0; j < DP; ++j) xm[j] = (j < D) ? (T)Xm[m * sXm0 + (int64_t)j * sXm1] : T(0);
  for (int k = 0; k < NK; ++k) {
    int64_t n = n0 + k;
    if (n >= N) break;
    T d[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) d[j] = xm[j] - xs[k][j];
    T q = T(0);
#pragma unroll
    for (int i = 0; i < DP; ++i) {
      if (i >= D) break;
      T y = T(0);
#pragma unroll
      for (int j = 0; j <= i; ++j) y = fma(Ls[k][i][j], d[j], y);
      q = fma(y, y, q);
    }
    out[n * so0 + m * so1] = (S)q;
  }
}

// D above 64: one thread per (n, m) pair, factor read from global memory.
template <typename T, typename S = T>
__global__ void __launch_bounds__(256)
quad_form_generic_kernel(const S *__restrict__ W, int64_t sW0, int64_t sW1, int64_t sW2,
                         const S *__restrict__ Xn, int64_t sXn0, int64_t sXn1,
                         const S *__restrict__ Xm, int64_t sXm0, int64_t sXm1,
                         S *__restrict__ out, int64_t so0, int64_t so1, int64_t N, int64_t M,
                         int64_t D) {
  int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t n = blockIdx.y;
  if (m >= M) return;
  T q = T(0);
  for (int64_t i = 0; i < D; ++i) {
    T y = T(0);
    for (int64_t j = 0; j <= i; ++j) {
      T dj = (T)Xm[m * sXm0 + j * sXm1] - (T)Xn[n * sXn0 + j * sXn1];
      y = fma((T)W[n * sW0 + j * sW1 + i * sW2], dj, y);
    }
    q = fma(y, y, q);
  }
  out[n * so0 + m * so1] = (S)q;
}

template <typename T, int DP, typename S = T>
static int launch_qf(const void *W, int64_t sW0, int64_t sW1, int64_t sW2, const void *Xn,
                     int64_t sXn0, int64_t sXn1, const void *Xm, int64_t sXm0, int64_t sXm1,
                     void *out, int64_t so0, int64_t so1, int64_t N, int64_t M, int64_t D,
                     hipStream_t s) {
  constexpr int NK = QfCfg<T, DP>::NK;
  dim3 grid((unsigned)ceil_div(M, 256), (unsigned)ceil_div(N, NK));
  hipLaunchKernelGGL((quad_form_kernel<T, DP, S>), grid, dim3(256), 0, s, (const S *)W, sW0, sW1,
                     sW2, (const S *)Xn, sXn0, sXn1, (const S *)Xm, sXm0, sXm1, (S *)out, so0,
                     so1, N, M, (int)D);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

template <typename T, typename S = T>
static int dispatch_qf(const void *W, int64_t sW0, int64_t sW1, int64_t sW2, const void *Xn,
                       int64_t sXn0, int64_t sXn1, const void *Xm, int64_t sXm0, int64_t sXm1,
                       void *out, int64_t so0, int64_t so1, int64_t N, int64_t M, int64_t D,
                       hipStream_t s) {
#define QF_ARGS W, sW0, sW1, sW2, Xn, sXn0, sXn1, Xm, sXm0, sXm1, out, so0, so1, N, M, D, s
  if (D <= 4) return launch_qf<T, 4, S>(QF_ARGS);
  if (D <= 8) return launch_qf<T, 8, S>(QF_ARGS);
  if (D <= 16) return launch_qf<T, 16, S>(QF_ARGS);
  if (D <= 32) return launch_qf<T, 32, S>(QF_ARGS);
  if (D <= 64) return launch_qf<T, 64, S>(QF_ARGS);
#undef QF_ARGS
  dim3 grid((unsigned)ceil_div(M, 256), (unsigned)N);
  hipLaunchKernelGGL((quad_form_generic_kernel<T, S>), grid, dim3(256), 0, s, (const S *)W, sW0,
                     sW1, sW2, (const S *)Xn, sXn0, sXn1, (const S *)Xm, sXm0, sXm1, (S *)out, so0,
                     so1, N, M, D);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// ------------------------------------------------------------- colsum / fill
__global__ void __launch_bounds__(256) fill_kernel(float *p, int64_t n, float v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void __launch_bounds__(256) colsum_kernel(const int32_t *__restrict__ col,
                                                    const float *__restrict__ val,
                                                    const int32_t *__restrict__ nnz_dev,
                                                    Fx *__restrict__ acc) {
  const int64_t nnz = *nnz_dev;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) fx_atomic_add(&acc[col[i]], val[i]);
}

// posterior contributions of one block (int32 CSR) fused with the start of
// a lazy-rebalance step (lakde_rebalance_lazy_begin): the true and stored
// column sums of the block's entries, and the unsafe flag cleared
int posterior_contribs_lazy(const int32_t *rowptr, const int32_t *col, const float *value,
                            const float *colscale, const float *Xn, const float *Xm, int64_t nb,
                            int64_t D, float nu0, float *S, float *nu, Fx *old_fx, Fx *U_old,
                            int32_t *unsafe, hipStream_t s) {
  LAKDE_CHECK_ARG(rowptr && col && value && colscale && Xn && Xm && S && nu && old_fx &&
                      U_old && unsafe && nb >= 1 && D >= 1 && D <= 512,
                  "bad args");
  return launch_outer<float, int32_t, int32_t, false>(rowptr, col, value, Xn, Xm, nb, D, 0, S, nu,
                                                      nu0, s, colscale,
                                                      OldSums{old_fx, U_old, unsafe});
}

}  // namespace lakde

using namespace lakde;

extern "C" {

int lakde_version(void) { return 1; }
const char *lakde_last_error(void) { return g_last_error.c_str(); }

int lakde_ind2ptr(const void *ind, int ind_dtype, int64_t nnz, int64_t M, void *out,
                  int out_dtype, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(M >= 0 && nnz >= 0, "negative size");
  LAKDE_CHECK_ARG(out != nullptr && (nnz == 0 || ind != nullptr), "null pointer");
  hipStream_t s = as_stream(stream);
  if (ind_dtype == LAKDE_I32 && out_dtype == LAKDE_I32) return launch_ind2ptr<int32_t, int32_t>(ind, nnz, M, out, s);
  if (ind_dtype == LAKDE_I32 && out_dtype == LAKDE_I64) return launch_ind2ptr<int32_t, int64_t>(ind, nnz, M, out, s);
  if (ind_dtype == LAKDE_I64 && out_dtype == LAKDE_I32) return launch_ind2ptr<int64_t, int32_t>(ind, nnz, M, out, s);
  if (ind_dtype == LAKDE_I64 && out_dtype == LAKDE_I64) return launch_ind2ptr<int64_t, int64_t>(ind, nnz, M, out, s);
  set_error("lakde_ind2ptr: index dtypes must be int32/int64");
  return LAKDE_EDTYPE;
}

int lakde_w_inv(const void *rowptr, int rowptr_dtype, const void *col, int col_dtype,
                const void *value, const void *Xn, const void *Xm, int64_t N, int64_t M,
                int64_t D, int fp_dtype, void *out, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 0 && M >= 0 && D >= 1, "bad size");
  LAKDE_CHECK_ARG(rowptr && col && value && Xn && Xm && out, "null pointer");
  LAKDE_CHECK_ARG(D <= 512, "D above 512 is not supported");
  (void)M;
  hipStream_t s = as_stream(stream);
  if (fp_dtype == LAKDE_F32)
    return dispatch_outer_idx<float>(rowptr_dtype, col_dtype, rowptr, col, value, Xn, Xm, N, D,
                                     out, nullptr, 0.f, s);
  if (fp_dtype == LAKDE_F64)
    return dispatch_outer_idx<double>(rowptr_dtype, col_dtype, rowptr, col, value, Xn, Xm, N, D,
                                      out, nullptr, 0.0, s);
  if (fp_dtype == LAKDE_F16)
    return dispatch_outer_idx<float, __half>(rowptr_dtype, col_dtype, rowptr, col, value, Xn, Xm,
                                             N, D, out, nullptr, 0.f, s);
  set_error("lakde_w_inv: value dtype must be float16/float32/float64");
  return LAKDE_EDTYPE;
}

int lakde_chol_quad_form(const void *W, int64_t sW0, int64_t sW1, int64_t sW2, const void *Xn,
                         int64_t sXn0, int64_t sXn1, const void *Xm, int64_t sXm0, int64_t sXm1,
                         void *out, int64_t so0, int64_t so1, int64_t N, int64_t M, int64_t D,
                         int fp_dtype, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 0 && M >= 0 && D >= 1, "bad size");
  LAKDE_CHECK_ARG(W && Xn && Xm && out, "null pointer");
  LAKDE_CHECK_ARG(N <= 65535LL * 4 || D <= 64, "too many kernels for one launch");
  if (N == 0 || M == 0) return LAKDE_OK;
  hipStream_t s = as_stream(stream);
  if (fp_dtype == LAKDE_F32)
    return dispatch_qf<float>(W, sW0, sW1, sW2, Xn, sXn0, sXn1, Xm, sXm0, sXm1, out, so0, so1, N,
                              M, D, s);
  if (fp_dtype == LAKDE_F64)
    return dispatch_qf<double>(W, sW0, sW1, sW2, Xn, sXn0, sXn1, Xm, sXm0, sXm1, out, so0, so1,
                               N, M, D, s);
  if (fp_dtype == LAKDE_F16)
    return dispatch_qf<float, __half>(W, sW0, sW1, sW2, Xn, sXn0, sXn1, Xm, sXm0, sXm1, out, so0,
                                      so1, N, M, D, s);
  set_error("lakde_chol_quad_form: dtype must be float16/float32/float64");
  return LAKDE_EDTYPE;
}

int64_t lakde_record_floats(int64_t D) { return rec_floats(D); }

int lakde_knn_prior_cov(const float *X, int64_t N, int64_t D, int64_t row0, int64_t nb,
                        const void *graph, int graph_dtype, int64_t K, float *out,
                        lakde_stream_t stream) {
  LAKDE_CHECK_ARG(X && graph && out, "null pointer");
  LAKDE_CHECK_ARG(D >= 1 && D <= 512 && K >= 1 && nb >= 0 && row0 >= 0 && row0 + nb <= N,
                  "bad size");
  hipStream_t s = as_stream(stream);
  const float *Xn = X + row0 * D;
  if (graph_dtype == LAKDE_I32)
    return launch_outer<float, int32_t, int32_t, true>(nullptr, graph, nullptr, Xn, X, nb, D, K,
                                                       out, nullptr, 0.f, s);
  if (graph_dtype == LAKDE_I64)
    return launch_outer<float, int32_t, int64_t, true>(nullptr, graph, nullptr, Xn, X, nb, D, K,
                                                       out, nullptr, 0.f, s);
  set_error("lakde_knn_prior_cov: graph dtype must be int32/int64");
  return LAKDE_EDTYPE;
}

int lakde_posterior_contribs(const void *rowptr, int rowptr_dtype, const void *col, int col_dtype,
                             const float *value, const float *colscale, const float *Xn,
                             const float *Xm, int64_t nb, int64_t M, int64_t D, float nu0,
                             float *S, float *nu, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(rowptr && col && value && Xn && Xm && S && nu, "null pointer");
  LAKDE_CHECK_ARG(nb >= 0 && M >= 0 && D >= 1 && D <= 512, "bad size");
  return dispatch_outer_idx<float>(rowptr_dtype, col_dtype, rowptr, col, value, Xn, Xm, nb, D, S,
                                   nu, nu0, as_stream(stream), colscale);
}

int lakde_fill_f32(float *p, int64_t n, float v, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(n >= 0 && (n == 0 || p), "bad args");
  if (n == 0) return LAKDE_OK;
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), p, n, v);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int lakde_colsum(const int32_t *col, const float *val, const int32_t *nnz_dev, int64_t N,
                 float *sums, lakde_fx128 *ws, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(col && val && nnz_dev && sums && ws && N >= 0, "bad args");
  if (N == 0) return LAKDE_OK;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(colsum_kernel, dim3(1024), dim3(256), 0, s, col, val, nnz_dev, ws);
  LAKDE_LAUNCHED();
  return colsum_finish(N, ws, sums, s);
}

}  // extern "C"
