// D-templated kernels of the fused E-step and the predictive density.  The
// launchers are explicitly instantiated per dimension in generated
// translation units (see lakde_amd/build.py) so they compile in parallel.
#pragma once

#include <cfloat>
#include <type_traits>

#include "qf.hpp"

namespace lakde {

// Device-side parameters of one E-step launch.
struct EParams {
  int64_t N, nb, row0, nchunks;
  const float *X;
  const float *rec;
  float threshold, log_t_lo;
  const float *c_old;
  const float *old_colsum;
  float *c;
  uint64_t *masks;   // nchunks x nb (chunk-major: a wave's 64 kernels are contiguous)
  int32_t *seg_tot;  // kESeg x nb: kept entries of kernel n in chunk segment s
  int32_t *row_tot;  // nb: kept entries of kernel n
                     // (both accumulated by the main kernel, cleared by estep_offsets_kernel)
  int32_t *seg_end;  // kESeg x nb: kept entries of kernel n in segments 0..s
  int64_t seg_len;   // chunks per segment
  int32_t *rowptr;   // nb + 1
  int32_t *out_row, *out_col;
  float *out_val;
  int64_t cap;
  int32_t *nnz_dev;
  const float *lazy_G;   // optional unsafe-step test of the lazy rebalance
  int32_t *lazy_unsafe;
  float *lazy_factor;    // optional: the factors of the columns to re-base
  // optional (compaction of a lazy-rebalance step): unless the step is
  // unsafe, store u = r / (G coeff) instead of r and accumulate the column
  // sums of r and u (the work of lazy_block_kernel)
  Fx *lazy_Rnew, *lazy_Unew;
  // optional (step driver): the row-pointer kernel also writes the entry
  // count and the status / unsafe flags into host-coherent memory as 64-bit
  // words (seq << 32 | value), each valid on its own when its tag is seq
  unsigned long long *flag_words;
  const int32_t *status_flag, *unsafe_flag;
  uint32_t flag_seq;
  // optional (with flag_words): 1 when the step needs the host (entries
  // beyond cap, status or unsafe set), else 0
  int32_t *abort_flag;
  // optional (UPDATE, D <= kSkipMaxD): workgroup g takes the columns
  // col_perm[64 g .. 64 g + 63] (a spatially compact chunk whose bounding
  // box is chunk_box[g] = lo[D], hi[D]) and skips every kernel whose log rho
  // is provably below (min lb of the chunk) - skip_margin on the whole box.
  // Kept pairs are then OR-ed into the masks at their own columns (the masks
  // must be zero on entry; the compaction clears what it reads) and counted
  // in their own segments one by one.
  const int32_t *col_perm;
  const float *chunk_box;
  float skip_margin;
};

constexpr int kSkipMaxD = 16;

__device__ __forceinline__ float clamp_nan(float x, float lo, float hi) {
  return (x != x) ? x : fminf(fmaxf(x, lo), hi);
}

// the window of G clamp(exp(c_old - c)) the stored form represents: stored
// values u = r / (G coeff) stay below 2^40 (r <= 1), inside the range of the
// fixed-point column sums, and far from float under/overflow
constexpr float kLazyGcLo = 0x1p-40f, kLazyGcHi = 0x1p40f;

// the lazy rebalance's unsafe-step test for column m of an UPDATE step
// (lakde_rebalance_lazy_check): G[m] clamp(exp(c_old - c)) outside the
// window or NaN; one atomic per wave with an unsafe column
// A column whose G coeff falls below the window (its normaliser rose by more
// than 40 ln 2 nats since its factor was last 1) is re-based when the
// factor and coefficient stay well inside float range: bit 1 of the flag,
// its factor in lazy_factor[m].  Anything else outside the window (NaN,
// above it, further below) takes the exact rebalance: bit 0.
constexpr float kLazyRebaseMin = 0x1p-100f;

__device__ __forceinline__ void lazy_unsafe_test(const EParams &p, bool valid, int64_t m,
                                                 float co, float cm) {
  bool bad = false, rebase = false;
  if (valid) {
    const float coeff = clamp_nan(expf(co - cm), FLT_MIN, FLT_MAX);
    const float gc = p.lazy_G[m] * coeff;
    if (!(gc >= kLazyGcLo && gc <= kLazyGcHi)) {
      rebase = p.lazy_factor != nullptr && gc < kLazyGcLo && gc >= kLazyRebaseMin &&
               coeff >= kLazyRebaseMin;
      bad = !rebase;
      if (rebase) p.lazy_factor[m] = gc;
    }
  }
  const bool any_bad = __ballot(bad) != 0, any_rebase = __ballot(rebase) != 0;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (any_bad) atomicOr(p.lazy_unsafe, 1);
    if (any_rebase) atomicOr(p.lazy_unsafe, 2);
  }
}

struct PParams {
  int64_t N, M, per_split, nsplit;
  const float *rec;
  const float *Y;
  float *part_m, *part_s;  // nsplit x M
  float *out;
};

constexpr int kEThreads = 256;  // 4 waves; lane = column, wave = kernel quarter
constexpr int kESeg = 256;      // chunk segments of the row scan and the compaction

// a written mask's kept entries, added to its (segment, kernel) count and
// to the kernel's row total
__device__ __forceinline__ void count_mask(const EParams &p, int64_t seg, int64_t n,
                                           uint64_t mask) {
  if (mask != 0ull) {
    const int32_t c = (int32_t)__popcll(mask);
    atomicAdd(&p.seg_tot[seg * p.nb + n], c);
    atomicAdd(&p.row_tot[n], c);
  }
}

// Upper bound of log rho_n(x) over the box [lo, hi] (box = lo[D], hi[D]):
// on the box each y_i = (Linv (x - x_n))_i lies within C_i +- R_i, with C_i
// the row applied to the box centre minus x_n and R_i = sum_j |L_ij| h_j
// (h = half-widths, inflated by 1e-6 against the rounding of centre and
// width), so |y_i| >= max(|C_i| - R_i, 0) and Q = |y|^2 has a lower bound
// LB; log rho = c + e Q with e = -nu/2 < 0.  LB is shrunk by 0.1% against
// the rounding of the sums; the callers' margins are tens of nats.  Two
// multiply-adds per factor entry.  r: the kernel's record (loaded per lane).
template <int D>
__device__ __forceinline__ float log_rho_box_ub(const float *__restrict__ r,
                                                const float *__restrict__ box) {
  float dm[D], hw[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    dm[j] = 0.5f * (box[j] + box[D + j]) - r[rec_x(D) + j];
    hw[j] = 0.5f * (box[D + j] - box[j]) * 1.000001f;
  }
  float lbq = 0.f;
  int t = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    float c = 0.f, rad = 0.f;
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      c = fmaf(r[t + j], dm[j], c);
      rad = fmaf(fabsf(r[t + j]), hw[j], rad);
    }
    t += i + 1;
    const float g = fmaxf(fabsf(c) - rad, 0.f);
    lbq = fmaf(g, g, lbq);
  }
  return r[rec_c(D)] + r[rec_c(D) + 1] * (0.999f * lbq);
}

// ---------------------------------------------------------------- E-step
// One workgroup per chunk of 64 columns (points m) and all nb kernels of the
// block.  Pass 1 (ACCUM/UPDATE): streaming log-sum-exp over the kernels,
// merged across the 4 waves in LDS, then the normaliser update.  Pass 2
// (UPDATE/FIXED): test every responsibility against the final normaliser
// and write the wave ballot as a 64-bit mask per (kernel, chunk).
//
// UPDATE mode does not recompute the quadratic forms in pass 2: the final
// normaliser c = logaddexp(lb, lse) is bounded below, already in pass 1, by
// B = max(lb, running max of the wave's log rho) (lb = logsubexp(c_old,
// log old_colsum + c_old) is known up front, the log-sum-exp is at least
// the largest term, and rounding is monotone), so a pair with
// fl(log rho - B) <= log_t_lo can never pass keep_resp.  The other pairs
// (typically a few per column and block) are queued per lane in LDS with
// their log rho and tested exactly in pass 2; a wave in which some lane
// queued more than kEQueue pairs falls back to recomputing all of its pairs.
constexpr int kEQueue = 8;

template <int D, int MODE, bool SKIP = false>
__global__ void __launch_bounds__(kEThreads)
estep_main_kernel(const float *__restrict__ rec, const float *__restrict__ X, EParams p) {
  __shared__ float sm_m[4][kWave], sm_s[4][kWave], sm_c[kWave];
  __shared__ int32_t q_n[MODE == LAKDE_ESTEP_UPDATE ? kEQueue : 1][kEThreads];
  __shared__ float q_x[MODE == LAKDE_ESTEP_UPDATE ? kEQueue : 1][kEThreads];
  __shared__ uint64_t sm_win[4][kWave];
  __shared__ int32_t sm_col[SKIP ? kWave : 1];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  // wave index made explicitly uniform so the record loads below are scalar
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int64_t chunk = blockIdx.x;
  // SKIP: the workgroup's columns are the chunk-th 64 of the order col_perm
  const int64_t slot = chunk * kWave + lane;
  const bool valid = slot < p.N;
  const int64_t m = SKIP ? (valid ? (int64_t)p.col_perm[slot] : p.N) : slot;
  if (SKIP && w == 0) sm_col[lane] = (int32_t)m;
  float xm[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xm[j] = valid ? X[m * D + j] : 0.f;
  const int64_t q = (p.nb + 3) / 4;
  const int64_t nbeg = w * q;
  const int64_t nend = (nbeg + q) < p.nb ? (nbeg + q) : p.nb;
  constexpr int64_t R = rec_floats(D);
  const int64_t diag_n = m - p.row0;
  float cm = 0.f;
  int cnt = 0;  // UPDATE: queued candidates of this lane (may exceed kEQueue)

  if (MODE != LAKDE_ESTEP_FIXED) {
    float lb = neg_inf();
    if (MODE == LAKDE_ESTEP_UPDATE && valid) {
      const float co = p.c_old[m];
      lb = logsubexp_f(co, logf(p.old_colsum[m]) + co);
    }
    // queue a pair whose exact log rho x can pass the keep test (B is a
    // lower bound of the final normaliser)
    auto cand = [&](float x, int64_t n, float B) {
      if (__fsub_rn(x, B) > p.log_t_lo) {
        if (cnt < kEQueue) {
          q_n[cnt][tid] = (int32_t)(n - nbeg);
          q_x[cnt][tid] = x;
        }
        ++cnt;
      }
    };
    LseE l;
    l.init();
    float B = lb;  // max(lb, running maximum): a lower bound of the final normaliser
    // only the waves whose columns overlap the block's own rows hold the
    // excluded diagonal pair (n == m - row0)
    const int64_t col0 = chunk * kWave - p.row0;
    const bool has_diag = SKIP ? __ballot(valid && diag_n >= nbeg && diag_n < nend) != 0ull
                               : (col0 < nend && col0 + kWave > nbeg);
    // four kernels n0..n3 (wave-uniform), then one kernel
    auto group4 = [&](auto diag_tag, int64_t n0, int64_t n1, int64_t n2, int64_t n3) {
      constexpr bool DIAG = decltype(diag_tag)::value;
      const float *r0 = rec + n0 * R, *r1 = rec + n1 * R;
      const float *r2 = rec + n2 * R, *r3 = rec + n3 * R;
      const float q0 = qf_rec<D>(r0, xm), q1 = qf_rec<D>(r1, xm);
      const float q2 = qf_rec<D>(r2, xm), q3 = qf_rec<D>(r3, xm);
      float x0 = log_rho_q<D>(r0, q0), x1 = log_rho_q<D>(r1, q1);
      float x2 = log_rho_q<D>(r2, q2), x3 = log_rho_q<D>(r3, q3);
      if (DIAG) {
        if (n0 == diag_n) x0 = neg_inf();
        if (n1 == diag_n) x1 = neg_inf();
        if (n2 == diag_n) x2 = neg_inf();
        if (n3 == diag_n) x3 = neg_inf();
      }
      const float g = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
      if (l.fold(g)) B = fmaxf(lb, g);
      l.s += (l.term(x0) + l.term(x1)) + (l.term(x2) + l.term(x3));
      if (MODE == LAKDE_ESTEP_UPDATE) {
        // one pre-test on the group's largest value (fl(x - B) is
        // monotone in x), then the group's pairs
        if (valid && __fsub_rn(g, B) > p.log_t_lo) {
          cand(x0, n0, B);
          cand(x1, n1, B);
          cand(x2, n2, B);
          cand(x3, n3, B);
        }
      }
    };
    auto one = [&](auto diag_tag, int64_t n) {
      constexpr bool DIAG = decltype(diag_tag)::value;
      float x = log_rho<D>(rec + n * R, xm);
      if (DIAG && n == diag_n) x = neg_inf();
      if (l.fold(x)) B = fmaxf(lb, x);
      l.s += l.term(x);
      if (MODE == LAKDE_ESTEP_UPDATE && valid) cand(x, n, B);
    };
    auto scan = [&](auto diag_tag) {
      if constexpr (SKIP) {
        // Kernels whose log rho stays below (the chunk's smallest lb) -
        // skip_margin on the chunk's box contribute less than
        // nb exp(-skip_margin) of the normaliser (below FP32 resolution) and
        // can never be kept: only the others are evaluated.  lb <= the final
        // normaliser, and -inf (no other mass) skips nothing.
        float lbmin = valid ? lb : __builtin_huge_valf();
#pragma unroll
        for (int off = kWave / 2; off >= 1; off >>= 1)
          lbmin = fminf(lbmin, __shfl_xor(lbmin, off, kWave));
        const float limit = lbmin - p.skip_margin;
        const float *box = p.chunk_box + chunk * 2 * D;
        for (int64_t wb = nbeg; wb < nend; wb += kWave) {
          const int64_t kn = wb + lane;
          const bool need = kn < nend && !(log_rho_box_ub<D>(rec + kn * R, box) < limit);
          uint64_t todo = __ballot(need);
          while (__popcll(todo) >= 4) {
            const int64_t n0 = wb + __builtin_ctzll(todo);
            todo &= todo - 1;
            const int64_t n1 = wb + __builtin_ctzll(todo);
            todo &= todo - 1;
            const int64_t n2 = wb + __builtin_ctzll(todo);
            todo &= todo - 1;
            const int64_t n3 = wb + __builtin_ctzll(todo);
            todo &= todo - 1;
            group4(diag_tag, n0, n1, n2, n3);
          }
          while (todo) {
            const int64_t n = wb + __builtin_ctzll(todo);
            todo &= todo - 1;
            one(diag_tag, n);
          }
        }
        return;
      }
      int64_t n = nbeg;
      if (D <= 24)
        for (; n + 4 <= nend; n += 4) group4(diag_tag, n, n + 1, n + 2, n + 3);
      for (; n < nend; ++n) one(diag_tag, n);
    };
    if (has_diag) scan(std::true_type{});
    else scan(std::false_type{});
    sm_m[w][lane] = l.m;
    sm_s[w][lane] = l.s;
    __syncthreads();
    if (w == 0) {
      Lse t;
      t.init();
#pragma unroll
      for (int ww = 0; ww < 4; ++ww) t.merge(sm_m[ww][lane], sm_s[ww][lane]);
      const float lse = t.value();
      if (valid) {
        if (MODE == LAKDE_ESTEP_ACCUM) cm = logaddexp_f(p.c[m], lse);
        else cm = logaddexp_f(lb, lse);
        p.c[m] = cm;
      }
      if (MODE == LAKDE_ESTEP_UPDATE && p.lazy_unsafe != nullptr)
        lazy_unsafe_test(p, valid, m, valid ? p.c_old[m] : 0.f, cm);
      sm_c[lane] = cm;
    }
    __syncthreads();
    cm = sm_c[lane];
  } else {
    cm = valid ? p.c[m] : 0.f;
  }
  if (MODE == LAKDE_ESTEP_ACCUM) return;
  const int64_t cseg = chunk / p.seg_len;

  if (MODE == LAKDE_ESTEP_UPDATE && __ballot(cnt > kEQueue) == 0) {
    // exact test of the queued candidates, 64 kernels at a time through a
    // per-wave LDS window of masks (LDS operations of one wave are in order)
    int ptr = 0;
    for (int64_t wb = 0; wb < nend - nbeg; wb += kWave) {
      sm_win[w][lane] = 0;
      __builtin_amdgcn_wave_barrier();
      while (ptr < cnt && q_n[ptr][tid] < wb + kWave) {
        const float d = __fsub_rn(q_x[ptr][tid], cm);
        float r;
        if (keep_resp(d, p.threshold, p.log_t_lo, &r))
          atomicOr((unsigned long long *)&sm_win[w][q_n[ptr][tid] - wb], 1ull << lane);
        ++ptr;
      }
      __builtin_amdgcn_wave_barrier();
      const int64_t n = nbeg + wb + lane;
      if (n < nend) {
        const uint64_t mk = sm_win[w][lane];
        if constexpr (SKIP) {
          // the chunk's columns are scattered: each kept pair sets its own
          // bit and counts in its own segment; the kernel's row total once
          // per wave
          if (mk != 0ull) {
            atomicAdd(&p.row_tot[n], (int32_t)__popcll(mk));
            uint64_t bits = mk;
            while (bits) {
              const int64_t mc = sm_col[__builtin_ctzll(bits)];
              bits &= bits - 1;
              atomicOr((unsigned long long *)&p.masks[(mc >> 6) * p.nb + n], 1ull << (mc & 63));
              atomicAdd(&p.seg_tot[((mc >> 6) / p.seg_len) * p.nb + n], 1);
            }
          }
        } else {
          p.masks[chunk * p.nb + n] = mk;
          count_mask(p, cseg, n, mk);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    return;
  }

  for (int64_t n = nbeg; n < nend; ++n) {
    float x = log_rho<D>(rec + n * R, xm);
    if (n == diag_n) x = neg_inf();
    const float d = __fsub_rn(x, cm);
    float r;
    const bool keep = valid && keep_resp(d, p.threshold, p.log_t_lo, &r);
    const uint64_t mask = __ballot(keep);
    if constexpr (SKIP) {
      if (keep) {
        atomicOr((unsigned long long *)&p.masks[(m >> 6) * p.nb + n], 1ull << (m & 63));
        atomicAdd(&p.seg_tot[((m >> 6) / p.seg_len) * p.nb + n], 1);
      }
      if (lane == 0 && mask != 0ull) atomicAdd(&p.row_tot[n], (int32_t)__popcll(mask));
    } else if (lane == 0) {
      p.masks[chunk * p.nb + n] = mask;
      count_mask(p, cseg, n, mask);
    }
  }
}

// Compaction: a wave per (64 kernels, chunk segment), lane = kernel n; the
// lane walks the segment's chunks in column order and writes each kept
// entry at rowptr[n] + seg_end[seg - 1][n] + (kept so far),
// recomputing it with the same device functions (identical bits) ->
// canonical CSR.  VALUES = false (large D): only the rows and columns are
// written here, and estep_values_kernel computes the values.
template <int D, bool VALUES = true>
__global__ void __launch_bounds__(256) estep_compact_kernel(EParams p) {
  const int64_t nnz = *p.nnz_dev;
  if (nnz > p.cap) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t n = (int64_t)blockIdx.x * kWave + lane;
  const int64_t seg = (int64_t)blockIdx.y * 4 + threadIdx.x / kWave;
  if (n >= p.nb || seg >= kESeg) return;
  const int64_t c0 = seg * p.seg_len;
  const int64_t c1 = (c0 + p.seg_len) < p.nchunks ? (c0 + p.seg_len) : p.nchunks;
  const int32_t s_lo = seg > 0 ? p.seg_end[(seg - 1) * p.nb + n] : 0;
  if (c0 >= c1 || p.seg_end[seg * p.nb + n] == s_lo) return;
  constexpr int64_t R = rec_floats(D);
  const float *rec = p.rec + n * R;
  int64_t pos = (int64_t)p.rowptr[n] + s_lo;
  const bool to_stored = p.lazy_Rnew != nullptr && (*p.lazy_unsafe & 1) == 0;
  // kept columns are gathered KB at a time so their point loads are in
  // flight together (the chain per lane is otherwise one load per entry)
  constexpr int KB = D <= 8 ? 4 : (D <= 24 ? 2 : 1);
  int64_t ms[KB];
  int k = 0;
  auto flush = [&]() {
    if constexpr (!VALUES) {
      for (int i = 0; i < k; ++i) {
        if (pos >= 0 && pos < p.cap) {  // never write outside the arena slots
          p.out_row[pos] = (int32_t)n;
          p.out_col[pos] = (int32_t)ms[i];
        }
        ++pos;
      }
      k = 0;
      return;
    }
    float xv[KB][D];
#pragma unroll
    for (int i = 0; i < KB; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) xv[i][j] = i < k ? p.X[ms[i] * D + j] : 0.f;
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      if (i < k) {
        const float x = log_rho<D>(rec, xv[i]);
        const float cm = p.c[ms[i]];
        const float d = __fsub_rn(x, cm);
        float r = 0.f;
        keep_resp(d, p.threshold, p.log_t_lo, &r);
        float v = r;
        if (to_stored) {
          const float coeff = clamp_nan(expf(p.c_old[ms[i]] - cm), FLT_MIN, FLT_MAX);
          v = r / (p.lazy_G[ms[i]] * coeff);
          fx_atomic_add(&p.lazy_Rnew[ms[i]], r);
          fx_atomic_add(&p.lazy_Unew[ms[i]], v);
        }
        if (pos >= 0 && pos < p.cap) {  // never write outside the arena slots
          p.out_row[pos] = (int32_t)n;
          p.out_col[pos] = (int32_t)ms[i];
          p.out_val[pos] = v;
        }
        ++pos;
      }
    }
    k = 0;
  };
  constexpr int MB = 16;  // mask loads in flight
  for (int64_t cb = c0; cb < c1; cb += MB) {
    uint64_t v[MB];
#pragma unroll
    for (int u = 0; u < MB; ++u) v[u] = (cb + u < c1) ? p.masks[(cb + u) * p.nb + n] : 0ull;
    {
      // the skipping and screening E-steps OR into the masks: leave them
      // zero (every nonzero mask lies in a segment with kept entries, so
      // the masks are all zero after every compaction)
#pragma unroll
      for (int u = 0; u < MB; ++u)
        if (v[u] != 0ull) p.masks[(cb + u) * p.nb + n] = 0ull;
    }
#pragma unroll
    for (int u = 0; u < MB; ++u) {
      uint64_t mask = v[u];
      while (mask) {
        ms[k++] = (cb + u) * kWave + __builtin_ctzll(mask);
        mask &= mask - 1;
        if (k == KB) flush();
      }
    }
  }
  if (k > 0) flush();
}

// Values of the compacted entries for large D: a wave per row n (the block's
// kernel), lanes = the row's entries, so the kernel's record is read once
// per wave (staged in LDS) instead of once per lane (the compaction's
// lanes are 64 different kernels).  Same device functions as the fused
// compaction: identical values, conversion and column sums.
template <int D>
__global__ void __launch_bounds__(256) estep_values_kernel(const float *__restrict__ recs,
                                                          const float *__restrict__ X,
                                                          EParams p) {
  const int64_t nnz = *p.nnz_dev;
  if (nnz > p.cap) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t n = (int64_t)blockIdx.x * 4 + w;
  if (n >= p.nb) return;
  constexpr int64_t R = rec_floats(D);  // a multiple of 4
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  // the wave's record staged in LDS with all its loads in flight, read back
  // as broadcast loads (a chain of dependent scalar loads otherwise)
  __shared__ f32x4 srec[4][R / 4];
  const bool to_stored = p.lazy_Rnew != nullptr && (*p.lazy_unsafe & 1) == 0;
  const int64_t beg = p.rowptr[n], end = p.rowptr[n + 1];
  if (beg >= end) return;
  const f32x4 *g = reinterpret_cast<const f32x4 *>(recs + n * R);
  constexpr int kPer = (int)((R / 4 + kWave - 1) / kWave);
  f32x4 st[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (lane + k * kWave < (int)(R / 4)) st[k] = g[lane + k * kWave];
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (lane + k * kWave < (int)(R / 4)) srec[w][lane + k * kWave] = st[k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float *rec = reinterpret_cast<const float *>(srec[w]);
  for (int64_t e0 = beg; e0 < end; e0 += kWave) {
    const int64_t e = e0 + lane;
    if (e < end) {
      const int64_t m = p.out_col[e];
      float xv[D];
#pragma unroll
      for (int j = 0; j < D; ++j) xv[j] = X[m * D + j];
      const float x = log_rho<D>(rec, xv);
      const float cm = p.c[m];
      const float d = __fsub_rn(x, cm);
      float r = 0.f;
      keep_resp(d, p.threshold, p.log_t_lo, &r);
      float v = r;
      if (to_stored) {
        const float coeff = clamp_nan(expf(p.c_old[m] - cm), FLT_MIN, FLT_MAX);
        v = r / (p.lazy_G[m] * coeff);
        fx_atomic_add(&p.lazy_Rnew[m], r);
        fx_atomic_add(&p.lazy_Unew[m], v);
      }
      p.out_val[e] = v;
    }
  }
}

// ------------------------------------------------------------ predictive
// grid (column chunks, kernel splits); each wave streams its share of the
// split's kernels for 64 test points; partial (max, sum) per split.
template <int D>
__global__ void __launch_bounds__(kEThreads)
pred_main_kernel(const float *__restrict__ rec, const float *__restrict__ Y, PParams p) {
  __shared__ float sm_m[4][kWave], sm_s[4][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t chunk = blockIdx.x;
  const int64_t split = blockIdx.y;
  const int64_t m = chunk * kWave + lane;
  const bool valid = m < p.M;
  float ym[D];
#pragma unroll
  for (int j = 0; j < D; ++j) ym[j] = valid ? Y[m * D + j] : 0.f;
  const int64_t sbeg = split * p.per_split;
  const int64_t send = (sbeg + p.per_split) < p.N ? (sbeg + p.per_split) : p.N;
  const int64_t q = (send - sbeg + 3) / 4;
  const int64_t nbeg = sbeg + w * q;
  const int64_t nend = (nbeg + q) < send ? (nbeg + q) : send;
  constexpr int64_t R = rec_floats(D);
  Lse l;
  l.init();
  int64_t n = nbeg;
  if (D <= 24) {
    for (; n + 4 <= nend; n += 4) {
      float x0 = log_pred<D>(rec + (n + 0) * R, ym);
      float x1 = log_pred<D>(rec + (n + 1) * R, ym);
      float x2 = log_pred<D>(rec + (n + 2) * R, ym);
      float x3 = log_pred<D>(rec + (n + 3) * R, ym);
      l.add4(x0, x1, x2, x3);
    }
  }
  for (; n < nend; ++n) l.add1(log_pred<D>(rec + n * R, ym));
  sm_m[w][lane] = l.m;
  sm_s[w][lane] = l.s;
  __syncthreads();
  if (w == 0 && valid) {
    Lse t;
    t.init();
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) t.merge(sm_m[ww][lane], sm_s[ww][lane]);
    p.part_m[split * p.M + m] = t.m;
    p.part_s[split * p.M + m] = t.s;
  }
}

// -------------------------------------------------- backward w.r.t. y
// d log p(y_m) / d y_m = sum_n w_nm (-(nu_n+1)/(1+Q_nm)) Linv_n^T z_nm with
// z = Linv (y - x_n) and w_nm = exp(lp_nm - LSE_m) recomputed from the saved
// per-point log-sum-exp; scaled by the upstream gradient g_m.  Same grid as
// the predictive kernel; per-split partial gradients are summed afterwards.
template <int D>
__global__ void __launch_bounds__(kEThreads)
bwd_y_kernel(const float *__restrict__ rec, const float *__restrict__ Y,
             const float *__restrict__ lse, const float *__restrict__ g, PParams p,
             float *__restrict__ part) {
  // one buffer, two uses: the lane's point y (y[j] at ys[j * 256]) during
  // the kernel loop, then the per-wave partial gradients for the fixed-order
  // combine.  y is read back per kernel instead of held in registers, which
  // leaves the difference, z and the gradient (3D floats) live per lane.
  __shared__ float sm_buf[4 * kWave * D];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t m = (int64_t)blockIdx.x * kWave + lane;
  const int64_t split = blockIdx.y;
  const bool valid = m < p.M;
  float *ys = sm_buf + threadIdx.x;
  float gy[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    ys[j * kEThreads] = valid ? Y[m * D + j] : 0.f;
    gy[j] = 0.f;
  }
  const float l_m = valid ? lse[m] : 0.f;
  const float g_m = valid ? g[m] : 0.f;
  const int64_t sbeg = split * p.per_split;
  const int64_t send = (sbeg + p.per_split) < p.N ? (sbeg + p.per_split) : p.N;
  const int64_t q = (send - sbeg + 3) / 4;
  const int64_t nbeg = sbeg + w * q;
  const int64_t nend = (nbeg + q) < send ? (nbeg + q) : send;
  constexpr int64_t R = rec_floats(D);
  for (int64_t n = nbeg; n < nend; ++n) {
    // compiler barrier: y is re-read from LDS for every kernel, not hoisted
    // into registers for the whole loop
    asm volatile("" ::: "memory");
    const float *r = rec + n * R;
    float d[D], z[D];
#pragma unroll
    for (int j = 0; j < D; ++j) d[j] = ys[j * kEThreads] - r[rec_x(D) + j];
    float Q = 0.f;
    int t = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float y = 0.f;
#pragma unroll
      for (int j = 0; j <= i; ++j) y = fmaf(r[t + j], d[j], y);
      t += i + 1;
      z[i] = y;
      Q = fmaf(y, y, Q);
    }
    const float h = r[rec_c(D) + 3];
    const float lp = __fsub_rn(r[rec_c(D) + 2], __fmul_rn(h, log1pf(Q)));
    const float wgt = expf(lp - l_m);
    const float a = -2.f * g_m * wgt * h / (1.f + Q);
    t = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const float az = a * z[i];
#pragma unroll
      for (int j = 0; j <= i; ++j) gy[j] = fmaf(r[t + j], az, gy[j]);
      t += i + 1;
    }
  }
  __syncthreads();  // every wave is done with its y before the buffer is reused
  float *gs = sm_buf + ((int64_t)w * kWave + lane) * D;
#pragma unroll
  for (int j = 0; j < D; ++j) gs[j] = gy[j];
  __syncthreads();
  if (w == 0 && valid) {
    const float *g0 = sm_buf + (int64_t)lane * D;
    constexpr int64_t WS = (int64_t)kWave * D;  // one wave's partials
#pragma unroll
    for (int j = 0; j < D; ++j)
      part[(split * p.M + m) * D + j] =
          (g0[j] + g0[WS + j]) + (g0[2 * WS + j] + g0[3 * WS + j]);
  }
}

// ------------------------------------------------------------- launchers
template <int D>
int launch_estep_main(const EParams &p, int mode, hipStream_t s) {
  dim3 grid((unsigned)p.nchunks);
  if (mode == LAKDE_ESTEP_ACCUM)
    hipLaunchKernelGGL((estep_main_kernel<D, LAKDE_ESTEP_ACCUM>), grid, dim3(kEThreads), 0, s, p.rec, p.X, p);
  else if (mode == LAKDE_ESTEP_FIXED)
    hipLaunchKernelGGL((estep_main_kernel<D, LAKDE_ESTEP_FIXED>), grid, dim3(kEThreads), 0, s, p.rec, p.X, p);
  else if (p.col_perm != nullptr) {
    if constexpr (D <= kSkipMaxD)
      hipLaunchKernelGGL((estep_main_kernel<D, LAKDE_ESTEP_UPDATE, true>), grid, dim3(kEThreads), 0,
                         s, p.rec, p.X, p);
    else
      return LAKDE_EINVAL;
  } else
    hipLaunchKernelGGL((estep_main_kernel<D, LAKDE_ESTEP_UPDATE>), grid, dim3(kEThreads), 0, s, p.rec, p.X, p);
  return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
}

// from this dimension on the values of the compacted entries are computed
// by estep_values_kernel (a wave per row) instead of in the mask walk
constexpr int kValuesMinD = 24;

template <int D>
int launch_estep_compact(const EParams &p, hipStream_t s) {
  if (p.nb == 0) return LAKDE_OK;
  const dim3 grid((unsigned)ceil_div(p.nb, kWave), (unsigned)(kESeg / 4));
  if constexpr (D >= kValuesMinD) {
    hipLaunchKernelGGL((estep_compact_kernel<D, false>), grid, dim3(256), 0, s, p);
    hipLaunchKernelGGL((estep_values_kernel<D>), dim3((unsigned)ceil_div(p.nb, 4)), dim3(256), 0,
                       s, p.rec, p.X, p);
  } else {
    hipLaunchKernelGGL((estep_compact_kernel<D, true>), grid, dim3(256), 0, s, p);
  }
  return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
}

template <int D>
int launch_pred_main(const PParams &p, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(p.M, kWave), (unsigned)p.nsplit);
  hipLaunchKernelGGL((pred_main_kernel<D>), grid, dim3(kEThreads), 0, s, p.rec, p.Y, p);
  return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
}

template <int D>
int launch_bwd_y(const PParams &p, const float *lse, const float *g, float *part,
                 hipStream_t s) {
  dim3 grid((unsigned)ceil_div(p.M, kWave), (unsigned)p.nsplit);
  hipLaunchKernelGGL((bwd_y_kernel<D>), grid, dim3(kEThreads), 0, s, p.rec, p.Y, lse, g, p, part);
  return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
}

// guard of the lazy rebalance kernels enqueued by the step driver before
// the host has read the step's flags (unsafe == nullptr: no guard)
struct LazyGuard {
  const int32_t *nnz;     // the block's entry count
  int64_t cap;
  const int32_t *unsafe;  // the unsafe-step flag
};

// shared host helpers (fused.hip)
int estep_compact_lazy(const lakde_estep_args &a, const float *G, int32_t *unsafe, Fx *Rnew,
                       Fx *Unew, hipStream_t s);
int estep_count(const lakde_estep_args &a, hipStream_t s,
                unsigned long long *flag_words = nullptr, const int32_t *status_flag = nullptr,
                const int32_t *unsafe_flag = nullptr, uint32_t flag_seq = 0,
                int32_t *abort_flag = nullptr);
int rebalance_lazy_guarded(int64_t N, const int32_t *col, float *val, const int32_t *nnz_dev,
                           const float *c_prev, float *c, float *G, Fx *U, Fx *U_old, Fx *ws,
                           LazyGuard guard, hipStream_t s);
int rebalance_lazy_rebase(int64_t N, int64_t num_blocks, int64_t cap, int64_t block,
                          const int32_t *nnz, const int32_t *cols, float *vals, float *G, Fx *U,
                          Fx *U_old, void *ws, const int32_t *unsafe, hipStream_t s);
float *rebase_factors(int64_t N, void *ws);
int rebalance_lazy_unsafe(int64_t N, int64_t num_blocks, int64_t cap, int64_t block,
                          const int32_t *nnz, const int32_t *cols, float *vals,
                          const float *c_prev, float *c, float *G, Fx *U, void *ws,
                          hipStream_t s);
void pred_geometry(int64_t N, int64_t M, int64_t *per_split, int64_t *nsplit);
int estep_fill_params(const lakde_estep_args &a, EParams *p);
int pred_fill_params(const float *rec, int64_t N, int64_t M, float *out, void *workspace,
                     size_t workspace_bytes, PParams *p);
int pred_combine(const PParams &p, lakde_stream_t stream);

// declarations for the dispatcher (defined by explicit instantiation)
#define LAKDE_DECL_D(d)                                                   \
  extern template int launch_estep_main<d>(const EParams &, int, hipStream_t); \
  extern template int launch_estep_compact<d>(const EParams &, hipStream_t);   \
  extern template int launch_pred_main<d>(const PParams &, hipStream_t);        \
  extern template int launch_bwd_y<d>(const PParams &, const float *, const float *, float *, \
                                      hipStream_t);

}  // namespace lakde
