// Entry points of the fused E-step, rebalance and predictive density, plus
// their dimension-independent kernels (row scans, rebalance passes, the
// split combine).  The D-templated kernels live in fused_impl.hpp.
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>

#include "fused_impl.hpp"

extern "C" size_t lakde_estep_workspace(int64_t N, int64_t nb);
extern "C" size_t lakde_pred_workspace(int64_t N, int64_t M);

namespace lakde {

LAKDE_FOR_EACH_D(LAKDE_DECL_D)

static int dispatch_estep_main(int64_t D, const EParams &p, int mode, hipStream_t s) {
  switch (D) {
#define CASE(d) case d: return launch_estep_main<d>(p, mode, s);
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  return LAKDE_EINVAL;
}

static int dispatch_estep_compact(int64_t D, const EParams &p, hipStream_t s) {
  switch (D) {
#define CASE(d) case d: return launch_estep_compact<d>(p, s);
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  return LAKDE_EINVAL;
}

static int dispatch_pred_main(int64_t D, const PParams &p, hipStream_t s) {
  switch (D) {
#define CASE(d) case d: return launch_pred_main<d>(p, s);
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  return LAKDE_EINVAL;
}

// ----------------------------------------------------------- row scans
// The main kernel adds every written mask's kept entries to its (segment,
// kernel) count and to the kernel's row total.  One launch then finishes the
// offsets: blocks 0 .. G-1 (16 waves x 64 kernels, a wave per 8 segments)
// turn the counts into inclusive per-segment ends (the compaction's write
// offsets within a row); block G scans the row totals into the row pointer
// and the entry count.  Both halves read only what the main kernel wrote,
// so the blocks need no handoff, and each clears its counters for the next
// E-step.
constexpr int kOffWaves = 16;
static_assert(kESeg % kOffWaves == 0, "segments per wave");

__global__ void __launch_bounds__(1024) estep_offsets_kernel(EParams p) {
  __shared__ int32_t part[kOffWaves][kWave];
  __shared__ int32_t sh[1024];
  const int tid = threadIdx.x;
  if (blockIdx.x + 1 < gridDim.x) {
    constexpr int SPW = kESeg / kOffWaves;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int64_t n = (int64_t)blockIdx.x * kWave + lane;
    const bool ok = n < p.nb;
    int32_t c[SPW];
#pragma unroll
    for (int k = 0; k < SPW; ++k) c[k] = ok ? p.seg_tot[(int64_t)(w * SPW + k) * p.nb + n] : 0;
    int32_t t = 0;
#pragma unroll
    for (int k = 0; k < SPW; ++k) t += c[k];
    part[w][lane] = t;
    __syncthreads();
    int32_t run = 0;
    for (int ww = 0; ww < w; ++ww) run += part[ww][lane];
    if (ok) {
#pragma unroll
      for (int k = 0; k < SPW; ++k) {
        run += c[k];
        p.seg_end[(int64_t)(w * SPW + k) * p.nb + n] = run;
        p.seg_tot[(int64_t)(w * SPW + k) * p.nb + n] = 0;
      }
    }
    return;
  }
  // last block: exclusive scan of the row totals
  const int64_t per = ceil_div(p.nb, 1024);
  const int64_t r0 = tid * per;
  const int64_t r1 = (r0 + per) < p.nb ? (r0 + per) : p.nb;
  int32_t s = 0;
  for (int64_t r = r0; r < r1; ++r) s += p.row_tot[r];
  sh[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t v = (tid >= off) ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += v;
    __syncthreads();
  }
  int32_t pos = sh[tid] - s;
  for (int64_t r = r0; r < r1; ++r) {
    p.rowptr[r] = pos;
    pos += p.row_tot[r];
    p.row_tot[r] = 0;
  }
  if (tid == 1023) {
    p.rowptr[p.nb] = sh[1023];
    *p.nnz_dev = sh[1023];
    if (p.abort_flag != nullptr)
      *p.abort_flag = (sh[1023] > p.cap || *p.status_flag != 0 || (*p.unsafe_flag & 1) != 0) ? 1 : 0;
    if (p.flag_words != nullptr) {
      const unsigned long long tag = (unsigned long long)p.flag_seq << 32;
      const unsigned long long w[3] = {tag | (uint32_t)sh[1023], tag | (uint32_t)*p.status_flag,
                                       tag | (uint32_t)*p.unsafe_flag};
#pragma unroll
      for (int i = 0; i < 3; ++i)
        __hip_atomic_store(&p.flag_words[i], w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// ------------------------------------------------------------ rebalance

// coeff[m] = clamp(exp(c_prev[m] - c[m]), tiny, max)
__global__ void __launch_bounds__(256) rebal_coeff_kernel(int64_t N, const float *c_prev,
                                                         const float *c, float *coeff) {
  int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  coeff[m] = clamp_nan(expf(c_prev[m] - c[m]), FLT_MIN, FLT_MAX);
}

// other blocks: v *= coeff[col], their old values summed in A[col]; the
// excluded block's values summed in B[col] (order-independent sums, so the
// column factor is applied to the sum, per column, afterwards).
// grid.y = block, grid.x strides over the block's nnz entries.
__global__ void __launch_bounds__(256) rebal_scale_sum_kernel(int64_t cap, int64_t exclude,
                                                             const int32_t *__restrict__ nnz,
                                                             const int32_t *__restrict__ cols,
                                                             float *__restrict__ vals,
                                                             const float *__restrict__ coeff,
                                                             Fx *__restrict__ A,
                                                             Fx *__restrict__ B) {
  const int64_t b = blockIdx.y;
  const int32_t n = nnz[b];
  const int32_t *cb = cols + b * cap;
  float *vb = vals + b * cap;
  const bool scale = b != exclude;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cb[e];
    const float v = vb[e];
    if (scale) {
      vb[e] = v * coeff[col];
      fx_atomic_add(&A[col], v);
    } else {
      fx_atomic_add(&B[col], v);
    }
  }
}

// s = clamp_min(coeff A + B, eps); c += log s; coeff = clamp(1/s, tiny, max);
// A, B cleared
__global__ void __launch_bounds__(256) rebal_norm_kernel(int64_t N, float *c, Fx *A, Fx *B,
                                                        float *coeff) {
  int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  float s = (float)((double)coeff[m] * fx_value(A[m]) + fx_value(B[m]));
  A[m] = fx_zero();
  B[m] = fx_zero();
  s = (s != s) ? s : fmaxf(s, FLT_EPSILON);
  c[m] = c[m] + logf(s);
  coeff[m] = clamp_nan(1.0f / s, FLT_MIN, FLT_MAX);
}

__global__ void __launch_bounds__(256) rebal_rescale_kernel(int64_t cap,
                                                           const int32_t *__restrict__ nnz,
                                                           const int32_t *__restrict__ cols,
                                                           float *__restrict__ vals,
                                                           const float *__restrict__ coeff) {
  const int64_t b = blockIdx.y;
  const int32_t n = nnz[b];
  const int32_t *cb = cols + b * cap;
  float *vb = vals + b * cap;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    vb[e] = vb[e] * coeff[cb[e]];
}

// A lazy step the stored form cannot represent, done exactly in two passes
// over the arena.  First: the other blocks' stored values made true (u G,
// relative to the old normalisers) and rescaled to the new ones, their
// sums in A; the step's own block, written as true values by the
// compaction, summed in B.  Second (after rebal_norm_kernel): every value
// times 1/s, and U = the new stored (= true, G = 1) column sums.  The same
// roundings as a fold, the exact rebalance and a second fold in sequence.
__global__ void __launch_bounds__(256) rebal_fold_scale_sum_kernel(
    int64_t cap, int64_t exclude, const int32_t *__restrict__ nnz, const int32_t *__restrict__ cols,
    float *__restrict__ vals, const float *__restrict__ G, const float *__restrict__ coeff,
    Fx *__restrict__ A, Fx *__restrict__ B) {
  const int64_t b = blockIdx.y;
  const int32_t n = nnz[b];
  const int32_t *cb = cols + b * cap;
  float *vb = vals + b * cap;
  const bool scale = b != exclude;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cb[e];
    if (scale) {
      const float v = vb[e] * G[col];
      vb[e] = v * coeff[col];
      fx_atomic_add(&A[col], v);
    } else {
      fx_atomic_add(&B[col], vb[e]);
    }
  }
}

__global__ void __launch_bounds__(256) rebal_rescale_sum_kernel(int64_t cap,
                                                               const int32_t *__restrict__ nnz,
                                                               const int32_t *__restrict__ cols,
                                                               float *__restrict__ vals,
                                                               const float *__restrict__ coeff,
                                                               Fx *__restrict__ U) {
  const int64_t b = blockIdx.y;
  const int32_t n = nnz[b];
  const int32_t *cb = cols + b * cap;
  float *vb = vals + b * cap;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cb[e];
    const float v = vb[e] * coeff[col];
    vb[e] = v;
    fx_atomic_add(&U[col], v);
  }
}

// Re-basing the columns whose normaliser rose past the stored form's window
// (lazy_unsafe_test: bit 1 of the flag, factor f = G coeff in the factor
// array): the other blocks' stored values of such a column become its true
// values relative to the step's normaliser (u f), their sum the column's
// stored total, and G = G / f, so that G coeff = 1 for the step.  The
// step's own block is skipped (its old entries are being replaced; the
// compaction writes the new ones with the new G).  Nothing happens for a
// step that takes the exact rebalance (bit 0), whose factors are cleared.
__global__ void __launch_bounds__(256) rebase_scan_kernel(int64_t num_blocks, int64_t cap,
                                                         int64_t exclude,
                                                         const int32_t *__restrict__ nnz,
                                                         const int32_t *__restrict__ cols,
                                                         float *__restrict__ vals,
                                                         const float *__restrict__ factor,
                                                         Fx *__restrict__ U2,
                                                         const int32_t *__restrict__ unsafe) {
  if (*unsafe != 2) return;  // re-basing only (a fixed grid: an idle step costs one read)
  for (int64_t b = blockIdx.x; b < num_blocks; b += gridDim.x) {
    if (b == exclude) continue;
    const int32_t n = nnz[b];
    const int32_t *cb = cols + b * cap;
    float *vb = vals + b * cap;
    for (int32_t e = threadIdx.x; e < n; e += blockDim.x) {
      const int32_t col = cb[e];
      const float f = factor[col];
      if (f != 0.f) {
        const float v = vb[e] * f;
        vb[e] = v;
        fx_atomic_add(&U2[col], v);
      }
    }
  }
}

__global__ void __launch_bounds__(256) rebase_finish_kernel(int64_t N, float *__restrict__ factor,
                                                           Fx *__restrict__ U2, float *__restrict__ G,
                                                           Fx *__restrict__ U, Fx *__restrict__ U_old,
                                                           const int32_t *__restrict__ unsafe) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  const float f = factor[m];
  if (f == 0.f) return;
  if (*unsafe == 2) {
    U[m] = U2[m];
    U_old[m] = fx_zero();
    G[m] = G[m] / f;
  }
  U2[m] = fx_zero();
  factor[m] = 0.f;
}

// ------------------------------------------------------- lazy rebalance
// Stored value u_e, true value u_e * G[col_e]; U[m] = sum of stored values
// of column m, an order-independent fixed-point sum: U - U_old is exact, so
// a column that empties reads exactly zero and is clamped to eps like the
// exact rebalance.

// before the E-step: true column sums of the block (for the normaliser
// update; finished into floats by colsum_finish_kernel) and its stored
// column sums
__global__ void __launch_bounds__(256) lazy_begin_kernel(const int32_t *__restrict__ cols,
                                                        const float *__restrict__ vals,
                                                        const int32_t *__restrict__ nnz_dev,
                                                        const float *__restrict__ G,
                                                        Fx *__restrict__ old_fx,
                                                        Fx *__restrict__ Uold,
                                                        int32_t *__restrict__ unsafe) {
  const int32_t n = *nnz_dev;
  if (blockIdx.x == 0 && threadIdx.x == 0) *unsafe = 0;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cols[e];
    const float u = vals[e];
    fx_atomic_add(&old_fx[col], u * G[col]);
    fx_atomic_add(&Uold[col], u);
  }
}

// sums[m] = float(acc[m]), acc cleared
__global__ void __launch_bounds__(256) colsum_finish_kernel(int64_t N, Fx *__restrict__ acc,
                                                           float *__restrict__ sums) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  sums[m] = (float)fx_value(acc[m]);
  acc[m] = fx_zero();
}

// after the E-step, before the host reads the entry count: flag columns
// where G coeff leaves [kLazyGcLo, kLazyGcHi] (or is NaN), which the stored
// representation does not cover; the caller then rebalances exactly
__global__ void __launch_bounds__(256) lazy_check_kernel(int64_t N, const float *__restrict__ c_prev,
                                                        const float *__restrict__ c,
                                                        const float *__restrict__ G,
                                                        int32_t *__restrict__ unsafe) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool bad = false;
  if (m < N) {
    const float gc = G[m] * clamp_nan(expf(c_prev[m] - c[m]), FLT_MIN, FLT_MAX);
    bad = !(gc >= kLazyGcLo && gc <= kLazyGcHi);
  }
  if (__ballot(bad) != 0 && (threadIdx.x & (kWave - 1)) == 0) atomicOr(unsafe, 1);
}

// A guard lets the step driver enqueue the lazy rebalance before it knows
// the step's outcome: the kernels do nothing when the unsafe flag is set or
// the block's entries did not fit the arena (the host then finishes the step).
__device__ __forceinline__ bool lazy_skip(const LazyGuard &g) {
  return g.unsafe != nullptr && ((*g.unsafe & 1) != 0 || (int64_t)*g.nnz > g.cap);
}

// after the E-step: the block's true values r -> stored u = r / (G coeff),
// with column sums of both
__global__ void __launch_bounds__(256) lazy_block_kernel(const int32_t *__restrict__ cols,
                                                        float *__restrict__ vals,
                                                        const int32_t *__restrict__ nnz_dev,
                                                        const float *__restrict__ c_prev,
                                                        const float *__restrict__ c,
                                                        const float *__restrict__ G,
                                                        Fx *__restrict__ Rnew,
                                                        Fx *__restrict__ Unew,
                                                        LazyGuard guard) {
  if (lazy_skip(guard)) return;
  const int32_t n = *nnz_dev;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cols[e];
    const float r = vals[e];
    const float coeff = clamp_nan(expf(c_prev[col] - c[col]), FLT_MIN, FLT_MAX);
    const float u = r / (G[col] * coeff);
    vals[e] = u;
    fx_atomic_add(&Rnew[col], r);
    fx_atomic_add(&Unew[col], u);
  }
}

// per column: s = max(coeff G (U - Uold) + Rnew, eps), c += log s,
// G *= coeff clamp(1/s), U += Unew - Uold (exactly); clears the per-step sums
__global__ void __launch_bounds__(256) lazy_norm_kernel(int64_t N, const float *__restrict__ c_prev,
                                                       float *__restrict__ c,
                                                       float *__restrict__ G,
                                                       Fx *__restrict__ U,
                                                       Fx *__restrict__ Uold,
                                                       Fx *__restrict__ Rnew,
                                                       Fx *__restrict__ Unew,
                                                       LazyGuard guard) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N || lazy_skip(guard)) return;
  const float cp = c_prev[m], cn = c[m];
  const float coeff = clamp_nan(expf(cp - cn), FLT_MIN, FLT_MAX);
  const float g = G[m];
  const Fx u = U[m], uo = Uold[m], un = Unew[m];
  const double um = fx_diff(u, uo);
  float s = (float)((double)(coeff * g) * um + fx_value(Rnew[m]));
  s = (s != s) ? s : fmaxf(s, FLT_EPSILON);
  c[m] = cn + logf(s);
  const float coeff2 = clamp_nan(1.0f / s, FLT_MIN, FLT_MAX);
  G[m] = g * coeff * coeff2;
  U[m] = fx_sub_add(u, uo, un);
  Uold[m] = fx_zero();
  Rnew[m] = fx_zero();
  Unew[m] = fx_zero();
}

// u *= G[col] over every block, U = exact stored column sums
__global__ void __launch_bounds__(256) fold_values_kernel(int64_t cap,
                                                         const int32_t *__restrict__ nnz,
                                                         const int32_t *__restrict__ cols,
                                                         float *__restrict__ vals,
                                                         const float *__restrict__ G,
                                                         Fx *__restrict__ U) {
  const int64_t b = blockIdx.y;
  const int32_t n = nnz[b];
  const int32_t *cb = cols + b * cap;
  float *vb = vals + b * cap;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int32_t col = cb[e];
    const float v = vb[e] * G[col];
    vb[e] = v;
    fx_atomic_add(&U[col], v);
  }
}

__global__ void __launch_bounds__(256) fold_reset_kernel(int64_t N, float *__restrict__ G) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  G[m] = 1.f;
}

__global__ void __launch_bounds__(256) zero_fx_kernel(Fx *p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = fx_zero();
}

// ----------------------------------------------------------- predictive
__global__ void __launch_bounds__(256) pred_combine_kernel(PParams p, float log_n) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= p.M) return;
  Lse t;
  t.init();
  for (int64_t s = 0; s < p.nsplit; ++s) t.merge(p.part_m[s * p.M + m], p.part_s[s * p.M + m]);
  p.out[m] = -log_n + t.value();
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

void pred_geometry(int64_t N, int64_t M, int64_t *per_split, int64_t *nsplit) {
  const int64_t chunks = ceil_div(M, kWave);
  int64_t ns = ceil_div(16384, chunks * 4);
  const int64_t max_ns = ceil_div(N, 256);
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  *per_split = ceil_div(N, ns);
  *nsplit = ceil_div(N, *per_split);
}

int estep_fill_params(const lakde_estep_args &a, EParams *p) {
  LAKDE_CHECK_ARG(a.D >= 1 && a.D <= 63, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(a.N >= 1 && a.nb >= 1 && a.row0 >= 0 && a.row0 + a.nb <= a.N, "bad block");
  LAKDE_CHECK_ARG(a.N < (1LL << 31), "N must fit int32 column indices");
  LAKDE_CHECK_ARG(a.X && a.rec && a.c, "null pointer");
  LAKDE_CHECK_ARG(a.mode == LAKDE_ESTEP_ACCUM || a.mode == LAKDE_ESTEP_FIXED ||
                      a.mode == LAKDE_ESTEP_UPDATE, "bad mode");
  if (a.mode == LAKDE_ESTEP_UPDATE) LAKDE_CHECK_ARG(a.c_old && a.old_colsum, "UPDATE needs c_old/old_colsum");
  if (a.mode != LAKDE_ESTEP_ACCUM) {
    LAKDE_CHECK_ARG(a.rowptr && a.nnz_dev && a.out_row && a.out_col && a.out_val, "null output");
    LAKDE_CHECK_ARG(a.workspace && a.workspace_bytes >= lakde_estep_workspace(a.N, a.nb),
                    "workspace too small");
    LAKDE_CHECK_ARG(a.threshold >= 0.f, "negative threshold");
  }
  p->N = a.N;
  p->nb = a.nb;
  p->row0 = a.row0;
  p->nchunks = ceil_div(a.N, kWave);
  p->X = a.X;
  p->rec = a.rec;
  p->threshold = a.threshold;
  // log-domain pre-test with a 1% margin: exp(d) > t is only evaluated when
  // d > log(t) - 0.01, which cannot drop an entry (expf error << 1%)
  p->log_t_lo = a.threshold > 0.f ? logf(a.threshold) - 0.01f : -110.f;
  p->c_old = a.c_old;
  p->old_colsum = a.old_colsum;
  p->c = a.c;
  char *ws = (char *)a.workspace;
  // The counters that must read zero on entry (seg_tot, row_tot) sit at the
  // END of the workspace, the scratch arrays at its start.  Blocks of
  // different sizes share one workspace (the last block of a fit is
  // smaller); with the counters anchored at the end, a launch with nb
  // kernels uses a part of the counter region of any larger block, and its
  // scratch lies inside the larger block's scratch, so the counters are never
  // overwritten and every launch finds (and leaves) them zero.
  const size_t segb = align256((size_t)kESeg * a.nb * 4);
  const size_t rowb = align256((size_t)a.nb * 4);
  // The masks start the workspace and the segment ends follow the masks of
  // the LARGEST block the workspace fits (nb_cap), so no block's segment
  // ends ever lie where another block's masks go: with the skipping layout
  // (col_perm) the masks must read zero on entry, and only the compaction
  // clears them.
  int64_t nb_cap = a.nb;
  if (ws != nullptr && a.workspace_bytes >= lakde_estep_workspace(a.N, a.nb)) {
    // lo fits; grow hi until it does not, then bisect
    int64_t lo = a.nb, hi = a.nb + 1;
    while (hi < (1LL << 30) && lakde_estep_workspace(a.N, hi) <= a.workspace_bytes) {
      lo = hi;
      hi *= 2;
    }
    while (hi - lo > 1) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (lakde_estep_workspace(a.N, mid) <= a.workspace_bytes) lo = mid;
      else hi = mid;
    }
    nb_cap = lo;
  }
  const size_t maskb = align256((size_t)nb_cap * p->nchunks * 8);
  char *end = ws ? (char *)(((uintptr_t)ws + a.workspace_bytes) & ~(uintptr_t)255) : ws + 512;
  p->seg_tot = (int32_t *)(end - 256 - segb - rowb);
  p->row_tot = (int32_t *)(end - 256 - rowb);
  p->masks = (uint64_t *)ws;
  p->seg_end = (int32_t *)(ws + maskb);
  p->seg_len = ceil_div(p->nchunks, kESeg);
  p->rowptr = a.rowptr;
  p->out_row = a.out_row;
  p->out_col = a.out_col;
  p->out_val = a.out_val;
  p->cap = a.cap;
  p->nnz_dev = a.nnz_dev;
  p->lazy_G = nullptr;
  p->lazy_unsafe = nullptr;
  p->lazy_factor = nullptr;
  p->flag_words = nullptr;
  p->status_flag = p->unsafe_flag = nullptr;
  p->flag_seq = 0;
  p->abort_flag = nullptr;
  p->lazy_Rnew = p->lazy_Unew = nullptr;
  p->col_perm = a.col_perm;
  p->chunk_box = a.chunk_box;
  // a skipped pair is at least skip_margin nats below its column's
  // normaliser: all of a block's skipped pairs together change the
  // normaliser by less than nb e^-30 (5e-11 relative for nb = 512, FP32
  // resolves 6e-8), and none can pass a threshold t (margin >= 8 - log t;
  // threshold 0 keeps r != 0, i.e. down to 2^-149 = e^-103.3); an ACCUM
  // pass keeps nothing, so only the normaliser bound applies
  const float norm_margin = 30.f + logf((float)a.nb / 512.f);
  p->skip_margin = a.mode == LAKDE_ESTEP_ACCUM ? norm_margin
                   : a.threshold > 0.f         ? fmaxf(norm_margin, 8.f - logf(a.threshold))
                                               : 110.f;
  if (a.col_perm != nullptr)
    LAKDE_CHECK_ARG(a.chunk_box != nullptr && a.D <= kSkipMaxD,
                    "col_perm needs chunk_box and D <= 16");
  if (a.mode == LAKDE_ESTEP_UPDATE && a.lazy_unsafe != nullptr) {
    LAKDE_CHECK_ARG(a.lazy_G != nullptr, "lazy_unsafe needs lazy_G");
    p->lazy_G = a.lazy_G;
    p->lazy_unsafe = a.lazy_unsafe;
    p->lazy_factor = a.lazy_factor;
  }
  return LAKDE_OK;
}


int pred_fill_params(const float *rec, int64_t N, int64_t M, float *out, void *workspace,
                     size_t workspace_bytes, PParams *p) {
  LAKDE_CHECK_ARG(N >= 1 && M >= 0, "bad size");
  LAKDE_CHECK_ARG(rec && out && workspace, "null pointer");
  LAKDE_CHECK_ARG(workspace_bytes >= lakde_pred_workspace(N, M), "workspace too small");
  p->N = N;
  p->M = M;
  pred_geometry(N, M, &p->per_split, &p->nsplit);
  p->rec = rec;
  p->Y = nullptr;
  p->part_m = (float *)workspace;
  p->part_s = (float *)((char *)workspace + align256((size_t)p->nsplit * M * 4));
  p->out = out;
  return LAKDE_OK;
}

int pred_combine(const PParams &p, lakde_stream_t stream) {
  hipLaunchKernelGGL(pred_combine_kernel, dim3((unsigned)ceil_div(p.M, 256)), dim3(256), 0,
                     as_stream(stream), p, (float)std::log((double)p.N));
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int rebalance_lazy_guarded(int64_t N, const int32_t *col, float *val, const int32_t *nnz_dev,
                           const float *c_prev, float *c, float *G, Fx *U, Fx *U_old, Fx *ws,
                           LazyGuard guard, hipStream_t s) {
  LAKDE_CHECK_ARG(N >= 1 && c_prev && c && G && U && U_old && ws, "bad args");
  LAKDE_CHECK_ARG((col && val && nnz_dev) || (!col && !val && !nnz_dev),
                  "col, val and nnz_dev must all be given or all be null");
  Fx *Rnew = ws, *Unew = ws + N;
  if (col) {
    hipLaunchKernelGGL(lazy_block_kernel, dim3(256), dim3(256), 0, s, col, val, nnz_dev, c_prev,
                       c, G, Rnew, Unew, guard);
    LAKDE_LAUNCHED();
  }
  hipLaunchKernelGGL(lazy_norm_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, N,
                     c_prev, c, G, U, U_old, Rnew, Unew, guard);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int colsum_finish(int64_t N, Fx *acc, float *sums, hipStream_t s) {
  hipLaunchKernelGGL(colsum_finish_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, N,
                     acc, sums);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// the compaction of a lazy-rebalance step, storing u = r / (G coeff) and
// accumulating the column sums of r and u unless the step is unsafe
// (lakde_rebalance_lazy's first half, fused)
int estep_compact_lazy(const lakde_estep_args &a, const float *G, int32_t *unsafe, Fx *Rnew,
                       Fx *Unew, hipStream_t s) {
  LAKDE_CHECK_ARG(a.mode == LAKDE_ESTEP_UPDATE && a.c_old && G && unsafe && Rnew && Unew,
                  "bad lazy compaction args");
  EParams p;
  int rc = estep_fill_params(a, &p);
  if (rc) return rc;
  p.lazy_G = G;
  p.lazy_unsafe = unsafe;
  p.lazy_Rnew = Rnew;
  p.lazy_Unew = Unew;
  rc = dispatch_estep_compact(a.D, p, s);
  if (rc) set_error("estep_compact_lazy: launch failed");
  return rc;
}

// the per-segment offsets, row pointer and entry count of an E-step (the
// first half of lakde_estep_finish; *nnz_dev is final afterwards)
int estep_count(const lakde_estep_args &a, hipStream_t s, unsigned long long *flag_words,
                const int32_t *status_flag, const int32_t *unsafe_flag, uint32_t flag_seq,
                int32_t *abort_flag) {
  EParams p;
  int rc = estep_fill_params(a, &p);
  if (rc) return rc;
  p.flag_words = flag_words;
  p.status_flag = status_flag;
  p.unsafe_flag = unsafe_flag;
  p.flag_seq = flag_seq;
  p.abort_flag = (flag_words != nullptr && status_flag != nullptr && unsafe_flag != nullptr)
                     ? abort_flag
                     : nullptr;
  hipLaunchKernelGGL(estep_offsets_kernel, dim3((unsigned)ceil_div(p.nb, kWave) + 1u), dim3(1024),
                     0, s, p);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// the exact rebalance's workspace (coefficients, then two zeroed fixed-point
// accumulators A and B) between exact steps: the factor array of the
// re-based columns in B's first N floats, their new totals in A; both are
// zero again after every step
float *rebase_factors(int64_t N, void *ws) {
  Fx *A = (Fx *)((char *)ws + ((size_t)N * 4 + 15) / 16 * 16);
  return reinterpret_cast<float *>(A + N);
}

int rebalance_lazy_rebase(int64_t N, int64_t num_blocks, int64_t cap, int64_t block,
                          const int32_t *nnz, const int32_t *cols, float *vals, float *G, Fx *U,
                          Fx *U_old, void *ws, const int32_t *unsafe, hipStream_t s) {
  Fx *A = (Fx *)((char *)ws + ((size_t)N * 4 + 15) / 16 * 16);
  float *factor = rebase_factors(N, ws);
  const unsigned grid = (unsigned)(num_blocks < 1024 ? num_blocks : 1024);
  hipLaunchKernelGGL(rebase_scan_kernel, dim3(grid), dim3(256), 0, s, num_blocks, cap, block, nnz,
                     cols, vals, factor, A, unsafe);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebase_finish_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, N,
                     factor, A, G, U, U_old, unsafe);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int rebalance_lazy_unsafe(int64_t N, int64_t num_blocks, int64_t cap, int64_t block,
                          const int32_t *nnz, const int32_t *cols, float *vals,
                          const float *c_prev, float *c, float *G, Fx *U, void *ws,
                          hipStream_t s) {
  LAKDE_CHECK_ARG(N >= 1 && num_blocks >= 1 && num_blocks <= 65535 && cap >= 1 &&
                      cap < (1LL << 31) && block >= 0 && block < num_blocks,
                  "bad size");
  LAKDE_CHECK_ARG(nnz && cols && vals && c_prev && c && G && U && ws, "null pointer");
  LAKDE_CHECK_ARG(((uintptr_t)ws & 15) == 0, "ws must be 16-byte aligned");
  float *coeff = (float *)ws;
  Fx *A = (Fx *)((char *)ws + ((size_t)N * 4 + 15) / 16 * 16), *B = A + N;
  const unsigned gN = (unsigned)ceil_div(N, 256);
  int64_t gx = ceil_div(cap, 256 * 4);
  if (gx > 64) gx = 64;
  const dim3 ge((unsigned)gx, (unsigned)num_blocks);
  hipLaunchKernelGGL(rebal_coeff_kernel, dim3(gN), dim3(256), 0, s, N, c_prev, c, coeff);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_fold_scale_sum_kernel, ge, dim3(256), 0, s, cap, block, nnz, cols, vals,
                     G, coeff, A, B);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_norm_kernel, dim3(gN), dim3(256), 0, s, N, c, A, B, coeff);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(zero_fx_kernel, dim3(gN), dim3(256), 0, s, U, N);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_rescale_sum_kernel, ge, dim3(256), 0, s, cap, nnz, cols, vals, coeff,
                     U);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(fold_reset_kernel, dim3(gN), dim3(256), 0, s, N, G);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

}  // namespace lakde

using namespace lakde;

extern "C" {

size_t lakde_estep_workspace(int64_t N, int64_t nb) {
  const int64_t nchunks = ceil_div(N, kWave);
  // scratch (masks, segment ends), the counters (segment counts, row
  // totals), 256 bytes at the end and 256 bytes of alignment slack
  return align256((size_t)nb * nchunks * 8) + 2 * align256((size_t)kESeg * nb * 4) +
         align256((size_t)nb * 4) + 256 + 256;
}

int lakde_estep_main(const lakde_estep_args *args, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr, "null args");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  rc = dispatch_estep_main(args->D, p, args->mode, as_stream(stream));
  if (rc) set_error("lakde_estep_main: launch failed");
  return rc;
}

int lakde_estep_finish(const lakde_estep_args *args, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr, "null args");
  if (args->mode == LAKDE_ESTEP_ACCUM) return LAKDE_OK;
  const int rc = estep_count(*args, as_stream(stream));
  if (rc) return rc;
  return lakde_estep_compact(args, stream);
}

int lakde_estep(const lakde_estep_args *args, lakde_stream_t stream) {
  int rc = lakde_estep_main(args, stream);
  if (rc) return rc;
  return lakde_estep_finish(args, stream);
}

int lakde_estep_compact(const lakde_estep_args *args, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr && args->mode != LAKDE_ESTEP_ACCUM, "bad args");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  rc = dispatch_estep_compact(args->D, p, as_stream(stream));
  if (rc) set_error("lakde_estep_compact: launch failed");
  return rc;
}

size_t lakde_rebalance_workspace(int64_t N) {
  return N < 0 ? 0 : ((size_t)N * 4 + 15) / 16 * 16 + 2 * (size_t)N * sizeof(Fx);
}

int lakde_rebalance(int64_t N, int64_t num_blocks, int64_t cap, int64_t exclude_block,
                    const int32_t *nnz, const int32_t *cols, float *vals, const float *c_prev,
                    float *c, void *ws, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 1 && num_blocks >= 1 && cap >= 1, "bad size");
  LAKDE_CHECK_ARG(nnz && cols && vals && c_prev && c && ws, "null pointer");
  LAKDE_CHECK_ARG(((uintptr_t)ws & 15) == 0, "ws must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  LAKDE_CHECK_ARG(num_blocks <= 65535 && cap < (1LL << 31), "too many blocks or entries");
  float *coeff = (float *)ws;
  Fx *A = (Fx *)((char *)ws + ((size_t)N * 4 + 15) / 16 * 16), *B = A + N;
  const unsigned gN = (unsigned)ceil_div(N, 256);
  // x-extent: enough workgroups per block to cover ~all entries in one pass
  int64_t gx = ceil_div(cap, 256 * 4);
  if (gx > 64) gx = 64;
  const dim3 ge((unsigned)gx, (unsigned)num_blocks);
  hipLaunchKernelGGL(rebal_coeff_kernel, dim3(gN), dim3(256), 0, s, N, c_prev, c, coeff);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_scale_sum_kernel, ge, dim3(256), 0, s, cap, exclude_block, nnz, cols,
                     vals, coeff, A, B);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_norm_kernel, dim3(gN), dim3(256), 0, s, N, c, A, B, coeff);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(rebal_rescale_kernel, ge, dim3(256), 0, s, cap, nnz, cols, vals, coeff);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int lakde_rebalance_lazy_begin(int64_t N, const int32_t *col, const float *val,
                               const int32_t *nnz_dev, const float *G, float *old_colsum,
                               lakde_fx128 *U_old, lakde_fx128 *ws, int32_t *unsafe,
                               lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 1 && col && val && nnz_dev && G && old_colsum && U_old && ws && unsafe,
                  "bad args");
  hipStream_t s = as_stream(stream);
  Fx *old_fx = ws + 2 * N;
  hipLaunchKernelGGL(lazy_begin_kernel, dim3(256), dim3(256), 0, s, col, val, nnz_dev, G, old_fx,
                     U_old, unsafe);
  LAKDE_LAUNCHED();
  return colsum_finish(N, old_fx, old_colsum, s);
}

int lakde_rebalance_lazy_check(int64_t N, const float *c_prev, const float *c, const float *G,
                               int32_t *unsafe, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 1 && c_prev && c && G && unsafe, "bad args");
  hipLaunchKernelGGL(lazy_check_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0,
                     as_stream(stream), N, c_prev, c, G, unsafe);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int lakde_rebalance_lazy(int64_t N, const int32_t *col, float *val, const int32_t *nnz_dev,
                         const float *c_prev, float *c, float *G, lakde_fx128 *U,
                         lakde_fx128 *U_old, lakde_fx128 *ws, lakde_stream_t stream) {
  return rebalance_lazy_guarded(N, col, val, nnz_dev, c_prev, c, G, U, U_old, ws,
                                LazyGuard{nullptr, 0, nullptr}, as_stream(stream));
}

int lakde_rebalance_fold(int64_t N, int64_t num_blocks, int64_t cap, const int32_t *nnz,
                         const int32_t *cols, float *vals, float *G, lakde_fx128 *U,
                         lakde_stream_t stream) {
  LAKDE_CHECK_ARG(N >= 1 && num_blocks >= 1 && num_blocks <= 65535 && cap >= 1, "bad size");
  LAKDE_CHECK_ARG(nnz && cols && vals && G && U, "null pointer");
  hipStream_t s = as_stream(stream);
  const unsigned gN = (unsigned)ceil_div(N, 256);
  hipLaunchKernelGGL(zero_fx_kernel, dim3(gN), dim3(256), 0, s, U, N);
  LAKDE_LAUNCHED();
  int64_t gx = ceil_div(cap, 256 * 4);
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(fold_values_kernel, dim3((unsigned)gx, (unsigned)num_blocks), dim3(256), 0, s,
                     cap, nnz, cols, vals, G, U);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(fold_reset_kernel, dim3(gN), dim3(256), 0, s, N, G);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int lakde_flag_words_alloc(void **out) {
  LAKDE_CHECK_ARG(out != nullptr, "null pointer");
  *out = nullptr;
  void *h = nullptr, *d = nullptr;
  LAKDE_HIP(hipHostMalloc(&h, 256, hipHostMallocCoherent | hipHostMallocMapped));
  std::memset(h, 0, 256);
  const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(h);
    LAKDE_HIP(e);
  }
  // word 4 keeps the device-side address of the words for the kernels
  reinterpret_cast<unsigned long long *>(h)[4] = (unsigned long long)(uintptr_t)d;
  *out = h;
  return LAKDE_OK;
}

int lakde_flag_words_free(void *words) {
  if (words != nullptr) LAKDE_HIP(hipHostFree(words));
  return LAKDE_OK;
}

size_t lakde_pred_workspace(int64_t N, int64_t M) {
  int64_t per, ns;
  pred_geometry(N, M, &per, &ns);
  return 2 * align256((size_t)ns * M * 4);
}

int lakde_pred_density(const float *rec, int64_t N, const float *Y, int64_t M, int64_t D,
                       float *out, void *workspace, size_t workspace_bytes,
                       lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 1 && D <= 63, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(Y != nullptr, "null pointer");
  PParams p;
  int rc = pred_fill_params(rec, N, M, out, workspace, workspace_bytes, &p);
  if (rc) return rc;
  if (M == 0) return LAKDE_OK;
  p.Y = Y;
  rc = dispatch_pred_main(D, p, as_stream(stream));
  if (rc) { set_error("lakde_pred_density: launch failed"); return rc; }
  return pred_combine(p, stream);
}

}  // extern "C"

// ------------------------------------------ predictive density, no allocation
// The per-split partial (max, sum-exp) arrays of lakde_pred_density live in
// this file-scope device array instead of a caller's workspace.  Each GPU
// holds its own copy for the life of the process; host code never uses the
// array's name as an address (on the host it is only a stand-in symbol) but
// the per-device address that hipGetSymbolAddress returns.
namespace lakde {

constexpr size_t kPredScratchBytes = LAKDE_PRED_NOSCRATCH_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[kPredScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// the current device's copy of `scratch` (looked up once per device)
static int pred_scratch(void **out) {
  int dev = 0;
  LAKDE_HIP(hipGetDevice(&dev));
  LAKDE_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    LAKDE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    LAKDE_CHECK_ARG(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LAKDE_OK;
}

}  // namespace lakde

extern "C" {

int lakde_pred_density_noscratch(const float *rec, int64_t N, const float *Y, int64_t M,
                                 int64_t D, float *out, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 1 && D <= 63, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(rec && Y && out && N >= 1 && M >= 0, "bad args");
  const size_t need = lakde_pred_workspace(N, M);
  if (need > kPredScratchBytes) {
    set_error("lakde_pred_density_noscratch: " + std::to_string(need) +
              " bytes of partial sums needed, the static scratch holds " +
              std::to_string(kPredScratchBytes) + " (use lakde_pred_density with a workspace)");
    return LAKDE_EINVAL;
  }
  if (M == 0) return LAKDE_OK;
  void *ws = nullptr;
  int rc = pred_scratch(&ws);
  if (rc) return rc;
  PParams p;
  rc = pred_fill_params(rec, N, M, out, ws, kPredScratchBytes, &p);
  if (rc) return rc;
  p.Y = Y;
  rc = dispatch_pred_main(D, p, as_stream(stream));
  if (rc) { set_error("lakde_pred_density_noscratch: launch failed"); return rc; }
  return pred_combine(p, stream);
}

}  // extern "C"
