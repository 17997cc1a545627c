// Precision form of the quadratic forms: y = Linv (x_m - x_n) on the gfx950
// BF16 matrix cores (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
//
// Split precision: both operands are pairs of bf16 values (hi = bf16(v),
// lo = bf16(v - hi)) and each K-step issues hi*hi + hi*lo + lo*hi (bf16x3,
// ~2^-17 relative per product).  The B operand is the difference
// x_m - x_n itself, formed in FP32 and split on the vector ALUs per kernel
// (hidden behind the matrix-core work), so there is no cancellation between
// large terms: the error is relative to |Linv (x_m - x_n)|, as in the FP32
// path.  The E-step kernels write the same outputs (normalisers, ballot
// masks) as the FP32 kernels; the compaction that follows recomputes every
// kept responsibility exactly in FP32 from the FP32 records, so stored
// values carry no bf16 error.
//
// Layouts (ISA of v_mfma_f32_32x32x16_bf16): A and B fragments hold 8 bf16 per
// lane, row/column = lane % 32, K = 16 * step + 8 * (lane / 32) + [0, 8);
// C holds 16 floats per lane, column = lane % 32, row = 8 * (r / 4) +
// 4 * (lane / 32) + r % 4 for register r.
#include "fused_impl.hpp"
#include "mfma32.hpp"
#include "screen.hpp"

namespace lakde {

LAKDE_FOR_EACH_D(LAKDE_DECL_SCREEN_D)
LAKDE_FOR_EACH_D(LAKDE_DECL_M32_D)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// bf16 hi/lo split of two floats (hi = bf16(d), lo = bf16(d - hi), both
// round-to-nearest-even), packed two per 32-bit word: the hi pair is
// converted once and widened back with a shift and a mask
__device__ __forceinline__ void split2(float d0, float d1, uint32_t &hi, uint32_t &lo) {
  const uint32_t hb = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){d0, d1}, bf16x2));
  const float h0 = __uint_as_float(hb << 16), h1 = __uint_as_float(hb & 0xffff0000u);
  hi = hb;
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){d0 - h0, d1 - h1}, bf16x2));
}

__host__ __device__ constexpr int bf_kp(int D) { return (D + 15) / 16 * 16; }
__host__ __device__ constexpr int bf_dp(int D) { return (D + 31) / 32 * 32; }
// per-kernel fragment block: L_hi [Dp x Kp] bf16, L_lo [Dp x Kp] bf16, x_n [Kp] f32
__host__ __device__ constexpr int64_t bf_frag_bytes(int D) {
  return (int64_t)bf_dp(D) * bf_kp(D) * 2 * 2 + (int64_t)bf_kp(D) * 4;
}

// ------------------------------------------------------------- packing
// Per kernel n of the records: dense Linv (zero padded) split into bf16
// hi/lo, and the centre x_n (zero padded) in FP32.
__global__ void __launch_bounds__(256) pack_frag_kernel(const float *__restrict__ rec, int64_t R,
                                                       int D, uint8_t *__restrict__ frag) {
  const int64_t n = blockIdx.x;
  const int Kp = bf_kp(D), Dp = bf_dp(D);
  const float *r = rec + n * R;
  uint8_t *f = frag + n * bf_frag_bytes(D);
  __bf16 *hi = (__bf16 *)f;
  __bf16 *lo = hi + Dp * Kp;
  float *xn = (float *)(lo + Dp * Kp);
  for (int t = threadIdx.x; t < Dp * Kp; t += blockDim.x) {
    const int i = t / Kp, j = t - i * Kp;
    const float v = (i < D && j <= i) ? r[i * (i + 1) / 2 + j] : 0.f;
    const __bf16 h = (__bf16)v;
    hi[t] = h;
    lo[t] = (__bf16)(v - (float)h);
  }
  for (int j = threadIdx.x; j < Kp; j += blockDim.x) xn[j] = j < D ? r[rec_x(D) + j] : 0.f;
}

// --------------------------------------------------- MFMA quadratic forms
// One wave owns 64 consecutive columns (two 32-column MFMA tiles); a
// workgroup of 4 waves owns 256 columns and walks the kernels in lockstep,
// double-buffering each kernel's fragment block through LDS.  After the
// MFMAs, lane l < 32 reduces tile 0's column l and lane l >= 32 tile 1's
// column l - 32, so every lane ends with Q for exactly one column (m =
// wave_base + lane) — the same lane -> column map as the FP32 kernels.
template <int RT, int KS>
struct BfTile {
  static constexpr int Kp = KS * 16;
  static constexpr int Dp = RT * 32;
  static constexpr int FB = Dp * Kp * 2 * 2 + Kp * 4;
  float xf[2][KS][8];  // this lane's point coordinates for its B fragments

  __device__ void load_points(const float *__restrict__ X, int64_t N, int D, int64_t base,
                              int lane) {
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int64_t col = base + ct * 32 + (lane & 31);
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = s * 16 + 8 * (lane >> 5) + e;
          xf[ct][s][e] = (col < N && k < D) ? X[col * D + k] : 0.f;
        }
    }
  }

  // hi/lo split of this lane's differences x_m - x_n for K-step s, both
  // column tiles
  __device__ __forceinline__ void split_step(const float *xn, int s, int lane, bf16x8 (&dh)[2],
                                             bf16x8 (&dl)[2]) const {
    float c[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) c[e] = xn[s * 16 + 8 * (lane >> 5) + e];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      u32x4 h, l;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        uint32_t hh, ll;
        split2(xf[ct][s][2 * e] - c[2 * e], xf[ct][s][2 * e + 1] - c[2 * e + 1], hh, ll);
        h[e] = hh;
        l[e] = ll;
      }
      dh[ct] = __builtin_bit_cast(bf16x8, h);
      dl[ct] = __builtin_bit_cast(bf16x8, l);
    }
  }

  // Q for this lane's column from one kernel's fragment block in LDS.
  // K-step by K-step: the split of step s + 1 and the epilogue of a row
  // tile whose last nonzero step is done are independent of step s's
  // MFMAs, so their vector instructions issue while the matrix core works.
  // Every accumulator sees the same MFMA sequence in the same order
  // (per step: hi*hi, hi*lo, lo*hi), and q sums row tiles in order.
  __device__ __forceinline__ float quad(const uint8_t *fb, int lane) const {
    const __bf16 *hi = reinterpret_cast<const __bf16 *>(fb);
    const __bf16 *lo = hi + Dp * Kp;
    const float *xn = reinterpret_cast<const float *>(lo + Dp * Kp);
    bf16x8 dh[KS][2], dl[KS][2];
    f32x16 acc[RT][2];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) acc[rt][0] = acc[rt][1] = (f32x16){};
    float q0 = 0.f, q1 = 0.f;
    split_step(xn, 0, lane, dh[0], dl[0]);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s + 1 < KS) split_step(xn, s + 1, lane, dh[s + 1], dl[s + 1]);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        // Linv is lower triangular: a K-step whose first column lies right
        // of the tile's last row multiplies only zeros
        if (s * 16 > rt * 32 + 31) continue;
        const int off = (rt * 32 + (lane & 31)) * Kp + s * 16 + 8 * (lane >> 5);
        const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(hi + off);
        const bf16x8 al = *reinterpret_cast<const bf16x8 *>(lo + off);
        acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, dh[s][0], acc[rt][0], 0, 0, 0);
        acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, dh[s][1], acc[rt][1], 0, 0, 0);
        acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, dl[s][0], acc[rt][0], 0, 0, 0);
        acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, dl[s][1], acc[rt][1], 0, 0, 0);
        acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, dh[s][0], acc[rt][0], 0, 0, 0);
        acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, dh[s][1], acc[rt][1], 0, 0, 0);
      }
      // row tile rt is complete after step min(KS - 1, 2 rt + 1); tiles
      // complete in increasing rt, so q sums them in order
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const int last = (2 * rt + 1 < KS - 1) ? 2 * rt + 1 : KS - 1;
        if (s != last) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          q0 = fmaf(acc[rt][0][r], acc[rt][0][r], q0);
          q1 = fmaf(acc[rt][1][r], acc[rt][1][r], q1);
        }
      }
    }
    const bool upper = lane >= 32;
    const float send = upper ? q0 : q1;
    const float recv = __shfl_xor(send, 32, kWave);
    return (upper ? q1 : q0) + recv;
  }
};

// Cooperative copy of one fragment block (FB bytes, 16-byte granules, 256
// threads) into LDS, split into its global loads, all issued together into
// registers, and its LDS stores.  The loops load the next kernel's block
// before the current kernel's matrix-core work and store it after, so the
// load latency is hidden behind the MFMAs instead of being waited for
// granule by granule.
template <int FB>
struct FragStage {
  static constexpr int kG = FB / 16, kFull = kG / 256, kTail = kG % 256;
  u32x4 r[kFull + 1];
  __device__ __forceinline__ void load(const uint8_t *__restrict__ src) {
    const u32x4 *s = reinterpret_cast<const u32x4 *>(src) + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kFull; ++i) r[i] = s[i * 256];
    if (kTail != 0 && threadIdx.x < kTail) r[kFull] = s[kFull * 256];
  }
  __device__ __forceinline__ void store(uint8_t *dst) const {
    u32x4 *d = reinterpret_cast<u32x4 *>(dst) + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kFull; ++i) d[i * 256] = r[i];
    if (kTail != 0 && threadIdx.x < kTail) d[kFull * 256] = r[kFull];
  }
};

template <int RT, int KS, int MODE>
__global__ void __launch_bounds__(256)
estep_bf16_kernel(const uint8_t *__restrict__ frag, const float *__restrict__ X,
                  const float *__restrict__ rec, int64_t R, int Dreal, EParams p) {
  using T = BfTile<RT, KS>;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2][T::FB];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t base = (int64_t)blockIdx.x * 256 + w * kWave;
  const int64_t m = base + lane;
  const bool valid = m < p.N;
  const int64_t chunk = base / kWave;
  T tile;
  tile.load_points(X, p.N, Dreal, base, lane);
  const int64_t diag_n = m - p.row0;
  const int64_t nb = p.nb;
  const int cofs = (int)rec_c(Dreal);
  float cm = 0.f;

  // UPDATE: pass 2 tests only the pairs queued in pass 1 (the bound of the
  // FP32 kernel: the normaliser is at least max(lb, running max))
  __shared__ int32_t q_n[MODE == LAKDE_ESTEP_UPDATE ? kEQueue : 1][256];
  __shared__ float q_x[MODE == LAKDE_ESTEP_UPDATE ? kEQueue : 1][256];
  __shared__ uint64_t sm_win[4][kWave];
  const int tid = threadIdx.x;
  int cnt = 0;

  if (MODE != LAKDE_ESTEP_FIXED) {
    float lb = neg_inf();
    if (MODE == LAKDE_ESTEP_UPDATE && valid) {
      const float co = p.c_old[m];
      lb = logsubexp_f(co, logf(p.old_colsum[m]) + co);
    }
    Lse l;
    l.init();
    FragStage<T::FB> st;
    st.load(frag);
    st.store(lds[0]);
    __syncthreads();
    for (int64_t n = 0; n < nb; ++n) {
      // the next block (the last one again on the last kernel) goes into the
      // other buffer, last read in iteration n - 1, which ended at a barrier
      st.load(frag + (n + 1 < nb ? n + 1 : n) * T::FB);
      const float q = tile.quad(lds[n & 1], lane);
      const float *rc = rec + n * R + cofs;
      float x = __fadd_rn(__fmul_rn(q, rc[1]), rc[0]);
      if (n == diag_n) x = neg_inf();
      l.add1(x);
      if (MODE == LAKDE_ESTEP_UPDATE && valid && __fsub_rn(x, fmaxf(lb, l.m)) > p.log_t_lo) {
        if (cnt < kEQueue) {
          q_n[cnt][tid] = (int32_t)n;
          q_x[cnt][tid] = x;
        }
        ++cnt;
      }
      st.store(lds[(n + 1) & 1]);
      __syncthreads();
    }
    const float lse = l.value();
    if (valid) {
      if (MODE == LAKDE_ESTEP_ACCUM) cm = logaddexp_f(p.c[m], lse);
      else cm = logaddexp_f(lb, lse);
      p.c[m] = cm;
    }
    if (MODE == LAKDE_ESTEP_UPDATE && p.lazy_unsafe != nullptr)
      lazy_unsafe_test(p, valid, m, valid ? p.c_old[m] : 0.f, cm);
  } else {
    cm = valid ? p.c[m] : 0.f;
  }
  if (MODE == LAKDE_ESTEP_ACCUM) return;
  const int64_t cseg = chunk / p.seg_len;

  // block-uniform decision (the recompute path below synchronises the block)
  const int ovf = MODE == LAKDE_ESTEP_UPDATE ? __syncthreads_or(cnt > kEQueue) : 1;
  if (!ovf) {
    int ptr = 0;
    for (int64_t wb = 0; wb < nb; wb += kWave) {
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
      const int64_t n = wb + lane;
      if (n < nb && chunk < p.nchunks) {
        const uint64_t mk = sm_win[w][lane];
        p.masks[chunk * p.nb + n] = mk;
        count_mask(p, cseg, n, mk);
      }
      __builtin_amdgcn_wave_barrier();
    }
    return;
  }

  FragStage<T::FB> st;
  st.load(frag);
  st.store(lds[0]);
  __syncthreads();
  for (int64_t n = 0; n < nb; ++n) {
    st.load(frag + (n + 1 < nb ? n + 1 : n) * T::FB);
    const float q = tile.quad(lds[n & 1], lane);
    const float *rc = rec + n * R + cofs;
    float x = __fadd_rn(__fmul_rn(q, rc[1]), rc[0]);
    if (n == diag_n) x = neg_inf();
    const float d = __fsub_rn(x, cm);
    float r;
    const bool keep = valid && keep_resp(d, p.threshold, p.log_t_lo, &r);
    const uint64_t mask = __ballot(keep);
    if (lane == 0 && chunk < p.nchunks) {
      p.masks[chunk * p.nb + n] = mask;
      count_mask(p, cseg, n, mask);
    }
    st.store(lds[(n + 1) & 1]);
    __syncthreads();
  }
}

template <int RT, int KS>
__global__ void __launch_bounds__(256)
pred_bf16_kernel(const uint8_t *__restrict__ frag, const float *__restrict__ Y,
                 const float *__restrict__ rec, int64_t R, int Dreal, PParams p) {
  using T = BfTile<RT, KS>;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2][T::FB];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t base = (int64_t)blockIdx.x * 256 + w * kWave;
  const int64_t m = base + lane;
  const int64_t split = blockIdx.y;
  T tile;
  tile.load_points(Y, p.M, Dreal, base, lane);
  const int64_t sbeg = split * p.per_split;
  const int64_t send = (sbeg + p.per_split) < p.N ? (sbeg + p.per_split) : p.N;
  const int cofs = (int)rec_c(Dreal);
  Lse l;
  l.init();
  FragStage<T::FB> st;
  st.load(frag + sbeg * T::FB);
  st.store(lds[0]);
  __syncthreads();
  for (int64_t n = sbeg; n < send; ++n) {
    const int64_t k = n - sbeg;
    st.load(frag + (n + 1 < send ? n + 1 : n) * T::FB);
    const float q = tile.quad(lds[k & 1], lane);
    const float *rc = rec + n * R + cofs;
    l.add1(__fsub_rn(rc[2], __fmul_rn(rc[3], log1pf(q))));
    st.store(lds[(k + 1) & 1]);
    __syncthreads();
  }
  if (m < p.M) {
    p.part_m[split * p.M + m] = l.m;
    p.part_s[split * p.M + m] = l.s;
  }
}

#define LAKDE_BF_SHAPES(X) X(1, 1) X(1, 2) X(2, 3) X(2, 4)

static bool bf_shape(int D, int *RT, int *KS) {
  if (D < 8 || D > 63) return false;
  *KS = bf_kp(D) / 16;
  *RT = bf_dp(D) / 32;
  // instantiated (RT, KS) pairs: D <= 16 -> (1,1); <= 32 -> (1,2); <= 48 -> (2,3); <= 64 -> (2,4)
  return (*RT == 1 && (*KS == 1 || *KS == 2)) || (*RT == 2 && (*KS == 3 || *KS == 4));
}

}  // namespace lakde

using namespace lakde;

extern "C" {

int64_t lakde_bf16_frag_bytes(int64_t D) { return (D >= 8 && D <= 63) ? bf_frag_bytes((int)D) : -1; }

int lakde_pack_frag_bf16(const float *rec, int64_t nb, int64_t D, void *frag,
                         lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 8 && D <= 63, "the bf16 form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(rec && frag && nb >= 0, "bad args");
  if (nb == 0) return LAKDE_OK;
  hipLaunchKernelGGL(pack_frag_kernel, dim3((unsigned)nb), dim3(256), 0, as_stream(stream), rec,
                     rec_floats(D), (int)D, (uint8_t *)frag);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

int lakde_estep_main_bf16(const lakde_estep_args *args, const void *frag,
                          lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr && frag, "null args");
  int RT, KS;
  LAKDE_CHECK_ARG(bf_shape((int)args->D, &RT, &KS), "the bf16 form needs 8 <= D <= 63");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  const dim3 grid((unsigned)ceil_div(p.N, 256));
  hipStream_t s = as_stream(stream);
  const int64_t R = rec_floats(args->D);
  const int Dr = (int)args->D;
#define LAUNCH(rt, ks)                                                                           \
  if (RT == rt && KS == ks) {                                                                    \
    if (args->mode == LAKDE_ESTEP_ACCUM)                                                         \
      hipLaunchKernelGGL((estep_bf16_kernel<rt, ks, LAKDE_ESTEP_ACCUM>), grid, dim3(256), 0, s,  \
                         (const uint8_t *)frag, p.X, p.rec, R, Dr, p);         \
    else if (args->mode == LAKDE_ESTEP_FIXED)                                                    \
      hipLaunchKernelGGL((estep_bf16_kernel<rt, ks, LAKDE_ESTEP_FIXED>), grid, dim3(256), 0, s,  \
                         (const uint8_t *)frag, p.X, p.rec, R, Dr, p);         \
    else                                                                                         \
      hipLaunchKernelGGL((estep_bf16_kernel<rt, ks, LAKDE_ESTEP_UPDATE>), grid, dim3(256), 0, s, \
                         (const uint8_t *)frag, p.X, p.rec, R, Dr, p);         \
  }
  LAKDE_BF_SHAPES(LAUNCH)
#undef LAUNCH
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

// ------------------------------------------------ screening E-step (screen.hpp)
int64_t lakde_screen_frag_bytes(int64_t D) {
  return (D >= 8 && D <= 63) ? sc_frag_bytes((int)D) : -1;
}

size_t lakde_screen_points_bytes(int64_t N, int64_t D) {
  return (D >= 8 && D <= 63 && N >= 0) ? (size_t)N * sc_kp((int)D) * 2 : 0;
}

int lakde_screen_points(const float *X, int64_t N, int64_t D, const float *origin, void *xb,
                        float *xnorm, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 8 && D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(X && origin && xb && xnorm && N >= 0, "bad args");
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d) case d: rc = launch_screen_points<d>(X, N, origin, (__bf16 *)xb, xnorm, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_screen_points: launch failed");
  return rc;
}

int lakde_pack_screen(const float *rec, int64_t nb, int64_t D, const float *origin, void *frag,
                      lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 8 && D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(rec && origin && frag && nb >= 0, "bad args");
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d) case d: rc = launch_pack_screen<d>(rec, nb, origin, (uint8_t *)frag, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_pack_screen: launch failed");
  return rc;
}

int lakde_screen_knn_bound(const float *rec, const float *X, int64_t N, int64_t D,
                           const void *graph, int graph_dtype, int64_t k, int64_t J, float *out,
                           lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 8 && D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(rec && X && graph && out && N >= 0 && k >= 0, "bad args");
  LAKDE_CHECK_ARG(graph_dtype == LAKDE_I32 || graph_dtype == LAKDE_I64, "graph must be int32/int64");
  LAKDE_CHECK_ARG(J >= 1 && J <= 64 && (J & (J - 1)) == 0, "J must be a power of two <= 64");
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d)                                                                               \
  case d: rc = launch_screen_knn_bound<d>(rec, X, N, graph, graph_dtype, k, (int)J, out, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_screen_knn_bound: launch failed");
  return rc;
}

// The split UPDATE form's ranges: about 128 kernels each, at most 8.
static void screen_split_geometry(int64_t nb, int *nsplit, int64_t *len) {
  int64_t sp = nb >= 256 ? nb / 128 : 1;
  if (sp > kMaxSplit) sp = kMaxSplit;
  *len = ceil_div(nb, sp);
  *nsplit = (int)ceil_div(nb, *len);
}

static size_t align256z(size_t x) { return (x + 255) / 256 * 256; }

size_t lakde_screen_split_workspace(int64_t N, int64_t nb) {
  if (N < 1 || nb < 1) return 0;
  int S;
  int64_t len;
  screen_split_geometry(nb, &S, &len);
  const size_t SN = (size_t)S * (size_t)N;
  return 2 * align256z(SN * 4) + align256z(SN) + align256z(SN * kKeepQ * 2) +
         align256z(SN * kKeepQ * 4) + align256z((size_t)N) + align256z(SN) +
         align256z(SN * kScreenQ * 4) + align256z((size_t)nb * 4) +
         2 * align256z((size_t)nb * kScreenListCap * 4) +
         align256z((size_t)nb * (kScreenListCap / 256) * 4) + 256;
}

int lakde_estep_main_screen_split(const lakde_estep_args *args, const void *frag, const void *xb,
                                  const float *xnorm, void *split_ws, size_t split_bytes,
                                  lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr && frag && xb && xnorm && split_ws, "null args");
  LAKDE_CHECK_ARG(args->D >= 8 && args->D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(args->nb <= 32767, "the screening E-step takes blocks of at most 32767 kernels");
  LAKDE_CHECK_ARG(args->col_perm == nullptr, "the screening E-step takes columns in index order");
  LAKDE_CHECK_ARG(split_bytes >= lakde_screen_split_workspace(args->N, args->nb),
                  "split workspace too small");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  SParams sp{(const uint8_t *)frag, (const __bf16 *)xb, xnorm, p.skip_margin, nullptr};
  if (args->mode == LAKDE_ESTEP_UPDATE) {
    screen_split_geometry(args->nb, &sp.nsplit, &sp.split_len);
    const size_t SN = (size_t)sp.nsplit * (size_t)args->N;
    char *w = (char *)split_ws;
    sp.part_m = (float *)w;
    w += align256z(SN * 4);
    sp.part_s = (float *)w;
    w += align256z(SN * 4);
    sp.kcount = (uint8_t *)w;
    w += align256z(SN);
    sp.kq_n = (int16_t *)w;
    w += align256z(SN * kKeepQ * 2);
    sp.kq_x = (float *)w;
    w += align256z(SN * kKeepQ * 4);
    sp.only = (uint8_t *)w;
    w += align256z((size_t)args->N);
    sp.sv_cnt = (uint8_t *)w;
    w += align256z(SN);
    sp.sv_e = (int32_t *)w;
    w += align256z(SN * kScreenQ * 4);
    sp.kl_cnt = (int32_t *)w;
    w += align256z((size_t)args->nb * 4);
    sp.kl_m = (int32_t *)w;
    w += align256z((size_t)args->nb * kScreenListCap * 4);
    sp.kl_x = (float *)w;
    w += align256z((size_t)args->nb * kScreenListCap * 4);
    sp.items = (int32_t *)w;
    w += align256z((size_t)args->nb * (kScreenListCap / 256) * 4);
    sp.items_total = (int32_t *)w;
    sp.kl_cap = kScreenListCap;
  }
  hipStream_t s = as_stream(stream);
  rc = LAKDE_EINVAL;
  switch (args->D) {
#define CASE(d) case d: rc = launch_estep_screen<d>(p, sp, args->mode, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_estep_main_screen_split: launch failed");
  return rc;
}

int lakde_estep_main_screen(const lakde_estep_args *args, const void *frag, const void *xb,
                            const float *xnorm, const float *lbound, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr && frag && xb && xnorm, "null args");
  LAKDE_CHECK_ARG(args->D >= 8 && args->D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(args->nb <= 32767, "the screening E-step takes blocks of at most 32767 kernels");
  LAKDE_CHECK_ARG(args->col_perm == nullptr, "the screening E-step takes columns in index order");
  LAKDE_CHECK_ARG(lbound == nullptr || args->mode == LAKDE_ESTEP_ACCUM,
                  "a normaliser bound is an input of ACCUM passes only");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  SParams sp{(const uint8_t *)frag, (const __bf16 *)xb, xnorm, p.skip_margin, lbound};
  hipStream_t s = as_stream(stream);
  rc = LAKDE_EINVAL;
  switch (args->D) {
#define CASE(d) case d: rc = launch_estep_screen<d>(p, sp, args->mode, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_estep_main_screen: launch failed");
  return rc;
}

// ------------------------------------------- FP32 matrix-core E-step (mfma32.hpp)
int64_t lakde_m32_frag_bytes(int64_t D) {
  return (D >= 17 && D <= 63) ? m32_frag_floats((int)D) * 4 : -1;
}

int lakde_pack_m32(const float *rec, int64_t nb, int64_t D, const float *origin, void *frag,
                   lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 17 && D <= 63, "the FP32 matrix-core E-step needs 17 <= D <= 63");
  LAKDE_CHECK_ARG(rec && origin && frag && nb >= 0, "bad args");
  LAKDE_CHECK_ARG(((uintptr_t)frag & 15) == 0, "frag must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d) case d: rc = launch_pack_m32<d>(rec, nb, origin, (float *)frag, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_pack_m32: launch failed");
  return rc;
}

int lakde_estep_main_m32(const lakde_estep_args *args, const void *frag, const float *xnorm,
                         lakde_stream_t stream) {
  LAKDE_CHECK_ARG(args != nullptr && frag && xnorm, "null args");
  LAKDE_CHECK_ARG(args->D >= 17 && args->D <= 63, "the FP32 matrix-core E-step needs 17 <= D <= 63");
  LAKDE_CHECK_ARG(args->nb <= 32767, "the FP32 matrix-core E-step takes blocks of at most 32767 kernels");
  LAKDE_CHECK_ARG(args->col_perm == nullptr, "the FP32 matrix-core E-step takes columns in index order");
  LAKDE_CHECK_ARG(((uintptr_t)frag & 15) == 0, "frag must be 16-byte aligned");
  EParams p;
  int rc = estep_fill_params(*args, &p);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  rc = LAKDE_EINVAL;
  switch (args->D) {
#define CASE(d) case d: rc = launch_estep_m32<d>(p, (const float *)frag, xnorm, args->mode, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) set_error("lakde_estep_main_m32: launch failed");
  return rc;
}

int lakde_pred_density_bf16(const float *rec, const void *frag, int64_t N, const float *Y,
                            int64_t M, int64_t D, float *out, void *workspace,
                            size_t workspace_bytes, lakde_stream_t stream) {
  int RT, KS;
  LAKDE_CHECK_ARG(bf_shape((int)D, &RT, &KS), "the bf16 form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(rec && frag && Y && out && workspace && N >= 1 && M >= 0, "bad args");
  PParams p;
  int rc = pred_fill_params(rec, N, M, out, workspace, workspace_bytes, &p);
  if (rc) return rc;
  if (M == 0) return LAKDE_OK;
  const dim3 grid((unsigned)ceil_div(M, 256), (unsigned)p.nsplit);
  hipStream_t s = as_stream(stream);
  const int64_t R = rec_floats(D);
#define LAUNCH(rt, ks)                                                                       \
  if (RT == rt && KS == ks)                                                                  \
    hipLaunchKernelGGL((pred_bf16_kernel<rt, ks>), grid, dim3(256), 0, s, (const uint8_t *)frag, \
                       Y, rec, R, (int)D, p);
  LAKDE_BF_SHAPES(LAUNCH)
#undef LAUNCH
  LAKDE_LAUNCHED();
  return pred_combine(p, stream);
}

// ------------------------------------ screening predictive density (screen.hpp)
// Kernel splits sized so that the launch is about eight waves of workgroups
// at three workgroups per CU (the per-split work then balances), with at
// least 512 kernels per split.
static void pred_screen_geometry(int64_t N, int64_t M, int64_t *per_split, int64_t *nsplit) {
  const int64_t groups = ceil_div(M > 0 ? M : 1, 256);
  int64_t ns = ceil_div(6144, groups);
  const int64_t max_ns = ceil_div(N, 512);
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  *per_split = ceil_div(N, ns);
  *nsplit = ceil_div(N, *per_split);
}

static size_t pred_screen_part_bytes(int64_t ns, int64_t M) {
  return ((size_t)ns * (size_t)M * 4 + 255) / 256 * 256;
}

size_t lakde_pred_screen_workspace(int64_t N, int64_t M) {
  if (N < 1 || M < 0) return 0;
  int64_t per, ns;
  pred_screen_geometry(N, M, &per, &ns);
  return 2 * pred_screen_part_bytes(ns, M);
}

int lakde_pred_density_screen(const float *rec, const void *frag, int64_t N, const float *Y,
                              int64_t M, int64_t D, const void *yb, const float *ynorm,
                              float *out, void *workspace, size_t workspace_bytes,
                              lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 8 && D <= 63, "the screening form needs 8 <= D <= 63");
  LAKDE_CHECK_ARG(rec && frag && Y && yb && ynorm && out && workspace && N >= 1 && M >= 0,
                  "bad args");
  LAKDE_CHECK_ARG(N < (1LL << 31), "at most 2^31 - 1 kernels");
  LAKDE_CHECK_ARG(workspace_bytes >= lakde_pred_screen_workspace(N, M), "workspace too small");
  LAKDE_CHECK_ARG(((uintptr_t)frag & 15) == 0, "frag must be 16-byte aligned");
  if (M == 0) return LAKDE_OK;
  PParams p;
  p.N = N;
  p.M = M;
  pred_screen_geometry(N, M, &p.per_split, &p.nsplit);
  p.rec = rec;
  p.Y = Y;
  p.part_m = (float *)workspace;
  p.part_s = (float *)((char *)workspace + pred_screen_part_bytes(p.nsplit, M));
  p.out = out;
  // all dropped terms of a split together stay below 1e-6 of its partial
  // sum (1e-6 nats): per_split e^-margin = 1e-6.  The Student-t terms fall
  // off only logarithmically in Q, so every nat of margin keeps many more
  // pairs in the exact phase than in the E-step
  const float margin = (float)(std::log((double)p.per_split) + std::log(1e6));
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d)                                                                                  \
  case d:                                                                                        \
    rc = launch_pred_screen<d>(p, (const uint8_t *)frag, (const __bf16 *)yb, ynorm, margin, s); \
    break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) {
    set_error("lakde_pred_density_screen: launch failed");
    return rc;
  }
  return pred_combine(p, stream);
}

}  // extern "C"
