// Precision form of the E-step (BF16 matrix cores + exact FP32 refinement).
//
// Every (kernel n, point m) pair of a block is first screened on the BF16
// matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulate) in the centre
// form
//     y~ = bf16(U_n) bf16(x_m - o) - c_n,   c_n = U_n (x_n - o) in FP32,
// with U_n = Linv_n, o a fixed origin (the data mean) and bf16(x_m - o)
// prepared once per fit.  When D is not a multiple of 16, -c_n enters as
// up to three more MFMA columns, bf16 parts of -c_n (head and rounding
// residues) against constant coordinates 1 of the points; otherwise it is
// the accumulators' initial value.
// Write v = fl(x_m - o) = (x_m - o)(1 + eps) (|eps| <= 2^-24), bf16(v) =
// v - r and bf16(U) = U - dU, with the residuals r and dU exact in FP32.  The
// error of y~ is -U r - dU v + dU r + U (x_m - o) eps, plus the FP32
// accumulation of the 65 terms of each row (<= 65 * 2^-24 (1 + u)^2
// sum_j |U_ij| |x_mj - o_j|, u = 2^-8) and the FP32 rounding of c_n, so by
// Cauchy-Schwarz
//     | y~ - U_n (x_m - x_n) |_2 <= E_nm = P_n r_m + Q_n a_m + C_n
// with F_n = |U_n|_F, dF_n = |dU_n|_F, r_m = |r|_2, a_m = |x_m - o|,
// b_n = |x_n - o| (all rounded up), P_n = F_n (1 + u) (U r and dU r, with
// dF_n <= u F_n), Q_n = dF_n (1 + 2^-24) + 67 * 2^-24 F_n (dU v, the
// accumulation and eps) and C_n = kb F_n b_n (kb = 1.2e-5 >= 65 * 2^-24
// (1 + u)^2, the rounding of c_n), plus the norm of what the bf16 parts
// leave of c_n, computed exactly when c_n is packed, in the form with the
// MFMA columns.  The
// actual residuals are about 0.4 u of
// their operands' norms on average, against the worst case u of the bound
// (2u + u^2) F_n a_m that the form without them would need.  Hence the exact
// log rho = const_n - nu_n/2 Q lies between
//     lo = const_n - nu_n/2 (sqrt Q~ + E)^2   and   hi = const_n - nu_n/2 max(sqrt Q~ - E, 0)^2
// (each widened by the FP32 rounding of Q~ and of the bound itself).
//
// A pair whose upper bound lies more than `margin` nats below B, a running
// lower bound of its column's new normaliser (max of the old-mass bound lb
// and of the pairs' lower bounds lo), contributes less than nb e^-margin of
// the normaliser and cannot be kept (margin = the pair-skipping E-step's,
// >= 30 nats, >= 8 - log threshold): it is dropped.  The others (about 0.5%
// of the pairs at D = 63, a few per column and block) are queued per lane in
// LDS; after the kernel loop each wave compacts its lanes' queues and
// recomputes the queued pairs exactly in FP32, one pair per lane, with
// log_rho<D> -- the function the compaction uses, so kept entries carry the
// same bits there.  The log-sum-exp, the normaliser update and the keep
// test then run on exact values only: the result differs from the FP32
// E-step only by the dropped pairs (< 5e-11 relative) and summation order.
//
// A lane whose queue fills up resumes in another pass over the kernels
// (block-uniform, rare); a lane with more keep candidates than its keep
// queue holds is finished by a keep pass (the FIXED-mode pass: threshold
// c_m + log t).  Kept pairs set their bits in the ballot masks with atomics
// (the masks read zero on entry; the compaction clears what it reads).
//
// MFMA layouts (v_mfma_f32_32x32x16_bf16): A and B fragments hold 8 bf16
// per lane, row/column = lane % 32, K = 16 s + 8 (lane / 32) + [0, 8); C
// holds 16 floats per lane, column = lane % 32, row = 8 (r / 4) +
// 4 (lane / 32) + r % 4 for register r.
#pragma once

#include "fused_impl.hpp"

namespace lakde {

typedef __bf16 sc_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sc_f32x16 __attribute__((ext_vector_type(16)));
typedef float sc_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t sc_u32x4 __attribute__((ext_vector_type(4)));

// error-bound coefficients (see above)
constexpr double kScreenU = 0x1p-8;     // bf16 unit roundoff
// FP32 accumulation of 65 terms (65 (1 + u)^2 * 2^-24) and fl(x - o) (2^-24)
constexpr double kScreenAcc = 67.0 * 0x1p-24;
constexpr float kScreenKb = 1.2e-5f;
constexpr int kScreenQ = 16;   // survivor queue entries per lane
constexpr int kKeepQ = 4;      // keep-candidate queue entries per lane
constexpr int kScreenMaxRange = 512;  // kernels per range with workgroup-aggregated list slots
// List slots per workgroup (LDS counters, one global atomic per kernel of the
// range) below D = 48: many survivors per kernel (mix64_d21: one contended
// global atomic per pair took 1.54 s per VI iteration against 0.44); from 48
// on the pairs survive sparsely and the aggregation's registers cost more
// than its atomics save (mix128_d63: 3.07 against 2.96 ms per E-step).
__host__ __device__ constexpr bool sc_agg_slots(int D) { return D < 48; }

__host__ __device__ constexpr int sc_kp(int D) { return (D + 15) / 16 * 16; }
__host__ __device__ constexpr int sc_dp(int D) { return (D + 31) / 32 * 32; }
// The operands hold the coordinates in reverse order (operand index
// j' = D - 1 - j), so that the factor U (lower triangular) becomes upper
// triangular: row i' has nonzeros in the columns j' >= i' only.  With
// padding columns (D < Kp) the centre term rides in the MFMA itself: the
// points carry the constant 1 in columns D .. D + sc_naug - 1 and the A
// tiles carry -c_n there, split into up to three bf16 parts (head, then the
// rounding residues), all in the last K-step that every row tile uses, so
// the accumulators start at zero and no tile is added.
__host__ __device__ constexpr int sc_naug(int D) {
  return sc_kp(D) - D < 3 ? sc_kp(D) - D : 3;
}
__host__ __device__ constexpr bool sc_aug(int D) { return sc_naug(D) > 0; }
// MFMA A tiles of U (32 rows x 16 columns) that are not all zero: K-step s
// of row tile rt holds nonzeros iff 16 s + 15 >= 32 rt; stored K-step-major
// in that order
__host__ __device__ constexpr bool sc_tile_used(int rt, int s) { return s * 16 + 15 >= rt * 32; }
__host__ __device__ constexpr int sc_ntiles(int D) {
  int t = 0;
  for (int s = 0; s < sc_kp(D) / 16; ++s)
    for (int rt = 0; rt < sc_dp(D) / 32; ++rt) t += sc_tile_used(rt, s) ? 1 : 0;
  return t;
}
__host__ __device__ constexpr int sc_tile_index(int D, int rt_, int s_) {
  int t = 0;
  for (int s = 0; s < sc_kp(D) / 16; ++s)
    for (int rt = 0; rt < sc_dp(D) / 32; ++rt) {
      if (s == s_ && rt == rt_) return t;
      t += sc_tile_used(rt, s) ? 1 : 0;
    }
  return -1;
}
// fragment block per kernel, laid out as its LDS image: the used A tiles of
// bf16(U) (1 KB each: lane l's 8 bf16 operand values at byte 16 l, row
// 32 rt + l % 32, columns 16 s + 8 (l / 32) + [0, 8), reversed coordinates),
// then scal[8]: P_n, Q_n, C_n (the error bound), the E-step constant and
// exponent (-nu/2) widened for the upper bound, then for the lower bound,
// one unused; then, without a padding column, -c_n [Dp] f32 indexed by
// operand row
constexpr int kScreenScal = 8;
__host__ __device__ constexpr int64_t sc_frag_bytes(int D) {
  return (int64_t)sc_ntiles(D) * 1024 + 4 * kScreenScal + (sc_aug(D) ? 0 : (int64_t)sc_dp(D) * 4);
}

// Parameters of the screening launches beyond EParams.
struct SParams {
  const uint8_t *frag;   // nb fragment blocks
  const __bf16 *xb;      // N x Kp: bf16(x_m - o), zero padded
  const float *xnorm;    // 2N: |x_m - o|, then the bf16 residual norms, rounded up
  float margin;          // screening margin (nats)
  const float *lbound;   // ACCUM (optional): a lower bound of each column's final normaliser
  // split UPDATE form (optional, see launch_estep_screen): the block's
  // kernels in nsplit ranges of split_len, one workgroup per (column group,
  // range); per (range, column) the partial log-sum-exp, the count of keep
  // candidates and the first kKeepQ of them
  int64_t split_len = 0;
  int nsplit = 0;
  float *part_m = nullptr, *part_s = nullptr;
  uint8_t *kcount = nullptr;
  int16_t *kq_n = nullptr;
  float *kq_x = nullptr;
  // FIXED (optional): only the columns with only[m] != 0 are processed
  uint8_t *only = nullptr;
  // split UPDATE form, survivors left to the later kernels: per (range,
  // column) sv_cnt[r N + m] of them (0 where the range finished them
  // itself), entry k at sv_e[(m nsplit + r) kScreenQ + k] = (slot + 1) << 15 |
  // kernel, slot = the pair's place in the kernel's list (-1: the list was
  // full); per kernel n of the block kl_cnt[n] pairs claimed, the columns of
  // the first kl_cap at kl_m[n kl_cap + slot] and their exact values at
  // kl_x[n kl_cap + slot]
  uint8_t *sv_cnt = nullptr;
  int32_t *sv_e = nullptr;
  int32_t *kl_cnt = nullptr;
  int32_t *kl_m = nullptr;
  float *kl_x = nullptr;
  int kl_cap = 0;
  // the nonempty (kernel << 8 | 256-entry chunk) work items of the lists,
  // items_total[0] of them
  int32_t *items = nullptr;
  int32_t *items_total = nullptr;
};

// ------------------------------------------------------------- preparation
// bf16(x_m - o) per point in reversed coordinates, the constant 1 in column
// D when there is a padding column, zero padded to Kp; xnorm[m] = |x_m - o|
// and xnorm[N + m] = |fl(x_m - o) - bf16(x_m - o)|, the norm of the
// rounding residual, both rounded up.
template <int D>
__global__ void __launch_bounds__(256) screen_points_kernel(const float *__restrict__ X, int64_t N,
                                                           const float *__restrict__ origin,
                                                           __bf16 *__restrict__ xb,
                                                           float *__restrict__ xnorm) {
  constexpr int Kp = sc_kp(D);
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= N) return;
  double s = 0.0, rs = 0.0;
#pragma unroll
  for (int j = 0; j < Kp; ++j) {
    float v = (j >= D && j < D + sc_naug(D)) ? 1.f : 0.f;  // the constant coordinates are exact
    if (j < D) {
      v = X[m * D + (D - 1 - j)] - origin[D - 1 - j];
      s += (double)v * (double)v;
    }
    const __bf16 vb = (__bf16)v;
    const double rj = (double)v - (double)(float)vb;
    rs += rj * rj;
    xb[m * Kp + j] = vb;
  }
  xnorm[m] = (float)(sqrt(s) * (1.0 + 1e-6));
  xnorm[N + m] = (float)(sqrt(rs) * (1.0 + 1e-6));
}

// Fragment blocks of nb kernels from their FP32 records: one wave per kernel.
template <int D>
__global__ void __launch_bounds__(64) pack_screen_kernel(const float *__restrict__ rec,
                                                         const float *__restrict__ origin,
                                                         uint8_t *__restrict__ frag) {
  constexpr int Kp = sc_kp(D), Dp = sc_dp(D);
  constexpr int64_t R = rec_floats(D);
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x;
  const float *r = rec + n * R;
  uint8_t *f = frag + n * sc_frag_bytes(D);
  constexpr int NT = sc_ntiles(D);
  __bf16 *uh = reinterpret_cast<__bf16 *>(f);
  float *scal = reinterpret_cast<float *>(f + NT * 1024);
  float *cneg = scal + kScreenScal;  // without a padding column only
  // c_n row by row (lane = row), |U|_F^2, |U - bf16(U)|_F^2 and |x_n - o|^2
  // in double
  double f2 = 0.0, d2 = 0.0, b2 = 0.0, e2 = 0.0;
  float c = 0.f;
  if (lane < D) {
    const float *row = r + lane * (lane + 1) / 2;
    for (int j = 0; j <= lane; ++j) {
      const float dj = r[rec_x(D) + j] - origin[j];
      c = fmaf(row[j], dj, c);
      f2 += (double)row[j] * (double)row[j];
      const double e = (double)row[j] - (double)(float)(__bf16)row[j];
      d2 += e * e;
    }
    const float dl = r[rec_x(D) + lane] - origin[lane];
    b2 = (double)dl * (double)dl;
    // what the bf16 parts of -c_n in the padding columns leave out
    float rest = -c;
    for (int k = 0; k < sc_naug(D); ++k) rest -= (float)(__bf16)rest;  // exact subtractions
    e2 = sc_aug(D) ? (double)rest * (double)rest : 0.0;
  }
  // operand row i' is row D - 1 - i' of U
  for (int s = 0; s < Kp / 16; ++s)
    for (int rt = 0; rt < Dp / 32; ++rt) {
      if (!sc_tile_used(rt, s)) continue;
      __bf16 *tile = uh + sc_tile_index(D, rt, s) * 512 + lane * 8;
      const int ip = rt * 32 + (lane & 31), i = D - 1 - ip;
      const float ci = __shfl(c, ip < D ? i : 0, kWave);
      for (int e = 0; e < 8; ++e) {
        const int jp = s * 16 + 8 * (lane >> 5) + e, j = D - 1 - jp;
        float v = 0.f;
        if (ip < D && jp < D && jp >= ip) v = r[i * (i + 1) / 2 + j];
        if (ip < D && jp >= D && jp < D + sc_naug(D)) {
          // part jp - D of -c_i: the bf16 head of what the earlier parts leave
          float rest = -ci;
          for (int k = 0; k < jp - D; ++k) rest -= (float)(__bf16)rest;
          v = rest;
        }
        tile[e] = (__bf16)v;
      }
    }
  if (!sc_aug(D) && lane < Dp) {
    const float ci = __shfl(c, lane < D ? D - 1 - lane : 0, kWave);
    cneg[lane] = lane < D ? -ci : 0.f;
  }
  f2 = wave_sum(f2);
  d2 = wave_sum(d2);
  b2 = wave_sum(b2);
  e2 = wave_sum(e2);
  if (lane == 0) {
    const double up = 1.0 + 1e-6;
    const double F = sqrt(f2) * up, dF = sqrt(d2) * up, b = sqrt(b2) * up;
    scal[0] = (float)(F * (1.0 + kScreenU) * up);
    scal[1] = (float)((dF * (1.0 + 0x1p-24) + kScreenAcc * F) * up);
    scal[2] = (float)(((double)kScreenKb * F * b + sqrt(e2)) * up * up);
    // log rho = cst + expo Q; hi = cst_hi + expo_hi Q_lo and lo = cst_lo +
    // expo_lo Q_hi in FP32, with the slack 0.01 + 1e-4 |cst| for the
    // rounding of the products, of the exact term and of these constants
    const double cst = r[rec_c(D) + 0], expo = r[rec_c(D) + 1];
    const double slack = 0.01 + 1e-4 * fabs(cst);
    scal[3] = (float)(cst + slack);
    scal[4] = (float)(expo * (1.0 - 1e-5));
    scal[5] = (float)(cst - slack);
    scal[6] = (float)(expo * (1.0 + 1e-5));
    scal[7] = 0.f;
  }
}

template <int D>
__device__ __forceinline__ float log_rho_lane(const float *rec, const float *__restrict__ xrow);

// A lower bound of every column's normaliser for the first log-sum-exp pass
// of a fit (ACCUM from -inf): the largest exact log rho over the kernels of
// the column's first J graph neighbours.  The log-sum-exp over all kernels
// is at least any of its terms (the maximum is not fooled by a repeated
// neighbour); the diagonal and out-of-range entries are skipped.  One thread
// per (point, neighbour), J (a power of two <= 64) consecutive lanes per
// point.
template <int D>
__global__ void __launch_bounds__(256) screen_knn_bound_kernel(
    const float *__restrict__ rec, const float *__restrict__ X, int64_t N,
    const void *__restrict__ graph, int gdtype, int64_t gk, int J, float *__restrict__ out) {
  constexpr int64_t R = rec_floats(D);
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t m = t / J;
  const int j = (int)(t - m * J);
  float v = neg_inf();
  if (m < N && j < gk) {
    const int64_t n = gdtype == LAKDE_I64 ? ((const int64_t *)graph)[m * gk + j]
                                          : (int64_t)((const int32_t *)graph)[m * gk + j];
    if (n >= 0 && n < N && n != m) v = log_rho_lane<D>(rec + n * R, X + m * D);
  }
  for (int d = J / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
  if (m < N && j == 0) out[m] = v;
}

// ---------------------------------------------------------------- kernel
// LDS stages of the kernel loop: kScreenKps consecutive fragment blocks
// (contiguous in global memory, each already in its LDS layout), two stages
// in a ring.  A stage is copied with direct global-to-LDS loads (16 bytes
// per lane, 1 KB per wave instruction, no registers): the copy of stage
// i + 1 is issued at the start of iteration i and has the whole iteration
// to land; the barrier at the end of the iteration waits for it
// (vmcnt(0)) and frees stage i for the copy after next.
constexpr int kScreenKps = 2;

template <int D>
struct ScreenLds {
  static constexpr int64_t kOne = sc_frag_bytes(D);  // one kernel, 16-byte multiple
  static constexpr int kBytes = (int)(kScreenKps * kOne);
  static constexpr int kInst = (kBytes + 1023) / 1024;  // wave instructions per full stage
};

typedef __attribute__((address_space(3))) void *sc_lds_ptr;

// issue the copy of `kernels` consecutive fragment blocks from src into the
// LDS stage dst (wave instruction j by wave j % 4; lane l moves the 16 bytes
// at 1024 j + 16 l, lanes past the end are inactive)
template <int D>
__device__ __forceinline__ void screen_stage_copy(const uint8_t *__restrict__ src, uint8_t *dst,
                                                  int kernels, int w, int lane) {
  using L = ScreenLds<D>;
  const int bytes = (int)(kernels * L::kOne);
#pragma unroll
  for (int i = 0; i < (L::kInst + 3) / 4; ++i) {
    const int j = w + 4 * i;  // wave-uniform
    const int off = j * 1024 + lane * 16;
    if (j * 1024 < bytes && off < bytes)
      __builtin_amdgcn_global_load_lds(src + off, (sc_lds_ptr)(dst + j * 1024), 16, 0, 0);
  }
}

// initial accumulators of one kernel's two column tiles: zero when the
// centre term rides in the padding column, else -c_n (cneg in the LDS
// image, indexed by operand row)
template <int D>
__device__ __forceinline__ void sc_init_acc(sc_f32x16 (&acc)[sc_dp(D) / 32][2],
                                            const float *cneg, int lane) {
#pragma unroll
  for (int rt = 0; rt < sc_dp(D) / 32; ++rt) {
    sc_f32x16 ci;
    if constexpr (sc_aug(D)) {
#pragma unroll
      for (int r = 0; r < 16; ++r) ci[r] = 0.f;
    } else {
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const sc_f32x4 v =
            *reinterpret_cast<const sc_f32x4 *>(cneg + rt * 32 + 8 * q4 + 4 * (lane >> 5));
        ci[4 * q4 + 0] = v[0];
        ci[4 * q4 + 1] = v[1];
        ci[4 * q4 + 2] = v[2];
        ci[4 * q4 + 3] = v[3];
      }
    }
    acc[rt][0] = ci;
    acc[rt][1] = ci;
  }
}

// per-wave flat pair list of the exact phase: entry j holds (kernel << 8 |
// lane) until lane j % 64 has read it, then the pair's exact log rho
struct ScreenWaveList {
  int32_t ent[kScreenQ * kWave];
};

// log_rho<D>(rec, xm) with a per-lane record: the same function, so the
// same bits; the record pointer is declared 16-byte aligned (records are),
// so that the fully unrolled loads of the packed triangle can be merged
// into 16-byte loads.
template <int D>
__device__ __forceinline__ float log_rho_lane(const float *rec, const float *__restrict__ xrow) {
  const float *r = reinterpret_cast<const float *>(__builtin_assume_aligned(rec, 16));
  float xv[D];
#pragma unroll
  for (int t = 0; t < D; ++t) xv[t] = xrow[t];
  return log_rho<D>(r, xv);
}

template <int D, int MODE, bool SPLIT = false>
__global__ void __launch_bounds__(256, 3)
estep_screen_kernel(EParams p, SParams sp) {
  using L = ScreenLds<D>;
  constexpr int Kp = sc_kp(D);
  constexpr int RT = sc_dp(D) / 32, KS = Kp / 16;
  constexpr int64_t R = rec_floats(D);
  constexpr int64_t FB = sc_frag_bytes(D);
  constexpr int kRing = 2 * L::kBytes > L::kBytes + (int)sizeof(ScreenWaveList[4])
                            ? 2 * L::kBytes
                            : L::kBytes + (int)sizeof(ScreenWaveList[4]);
  // the two stages of the kernel loop (two arrays, so that the compiler can
  // tell the copy into one from the reads of the other); the exact phase
  // after the loop reuses the memory for the waves' pair lists
  __shared__ __attribute__((aligned(1024))) uint8_t ring0[kRing - L::kBytes];
  __shared__ __attribute__((aligned(1024))) uint8_t ring1[L::kBytes];
  __shared__ int16_t sq[kScreenQ][256];
  ScreenWaveList *wl = reinterpret_cast<ScreenWaveList *>(ring0);
  // the split form's list-slot aggregation (per-kernel counters and bases,
  // the lanes' local slots): in the stage arrays when they are large enough
  constexpr int kAggCnt = 2 * kScreenMaxRange * 4, kAggSlots = kScreenQ * 256 * 2;
  constexpr bool kAggRing = kRing - L::kBytes >= kAggCnt && L::kBytes >= kAggSlots;
  __shared__ __attribute__((aligned(16)))
  uint8_t agg_own[(SPLIT && sc_agg_slots(D) && !kAggRing) ? kAggCnt + kAggSlots : 16];
  __shared__ int16_t kq_n[MODE == LAKDE_ESTEP_UPDATE ? kKeepQ : 1][256];
  __shared__ float kq_x[MODE == LAKDE_ESTEP_UPDATE ? kKeepQ : 1][256];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int64_t base = (int64_t)blockIdx.x * 256 + w * kWave;  // the wave's 64 columns = one chunk
  const int64_t m = base + lane;
  // FIXED with a column filter: the other columns take no part
  const bool valid = m < p.N && (MODE != LAKDE_ESTEP_FIXED || sp.only == nullptr || sp.only[m] != 0);
  if (MODE == LAKDE_ESTEP_FIXED && sp.only != nullptr && !__syncthreads_or(valid)) return;
  const int64_t chunk = base / kWave;
  const int64_t nb = p.nb;
  // this workgroup's kernels [kb_, ke_) (SPLIT: one range of the block)
  const int64_t kb_ = SPLIT ? (int64_t)blockIdx.y * sp.split_len : 0;
  const int64_t ke_ = SPLIT ? (kb_ + sp.split_len < nb ? kb_ + sp.split_len : nb) : nb;
  const int64_t diag_n = m - p.row0;

  // this lane's B fragments (both column tiles) and column norm
  sc_bf16x8 xf[2][KS];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    const int64_t col = base + ct * 32 + (lane & 31);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (col < p.N)
        xf[ct][s] = *reinterpret_cast<const sc_bf16x8 *>(sp.xb + col * Kp + s * 16 + 8 * (lane >> 5));
      else
        xf[ct][s] = __builtin_bit_cast(sc_bf16x8, (sc_u32x4){0u, 0u, 0u, 0u});
    }
  }
  const float am = valid ? sp.xnorm[m] : 0.f;
  const float rm = valid ? sp.xnorm[p.N + m] : 0.f;

  float lb = neg_inf();   // UPDATE: bound of the normaliser from the old mass
  float cm = 0.f;         // ACCUM, FIXED: the given normaliser; later the new one
  if (MODE == LAKDE_ESTEP_UPDATE && valid) {
    const float co = p.c_old[m];
    lb = logsubexp_f(co, logf(p.old_colsum[m]) + co);
  }
  if (MODE != LAKDE_ESTEP_UPDATE && valid) cm = p.c[m];
  // running lower bound of the new normaliser (ACCUM: the new normaliser
  // logaddexp(c, LSE) is at least c, and the final one at least the given
  // bound)
  float B = MODE == LAKDE_ESTEP_ACCUM ? cm : lb;
  if (MODE == LAKDE_ESTEP_ACCUM && sp.lbound != nullptr && valid) B = fmaxf(B, sp.lbound[m]);
  Lse lse;
  lse.init();
  int kcnt = 0;           // UPDATE: keep candidates queued (may exceed kKeepQ)

  // bounds [lo, hi] of the exact log rho of kernel n (LDS image fb) at this
  // lane's column; -inf for the excluded diagonal and invalid columns
  auto bounds = [&](const uint8_t *fb, int64_t n, float &hi, float &lo) {
    const float *scal = reinterpret_cast<const float *>(fb + sc_ntiles(D) * 1024);
    sc_f32x16 acc[RT][2];
    sc_init_acc<D>(acc, scal + kScreenScal, lane);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        if (!sc_tile_used(rt, s)) continue;  // upper triangular in operand order
        const sc_bf16x8 a = *reinterpret_cast<const sc_bf16x8 *>(
            fb + sc_tile_index(D, rt, s) * 1024 + lane * 16);
        acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xf[0][s], acc[rt][0], 0, 0, 0);
        acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xf[1][s], acc[rt][1], 0, 0, 0);
      }
    }
    float q0 = 0.f, q1 = 0.f;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        q0 = fmaf(acc[rt][0][r], acc[rt][0][r], q0);
        q1 = fmaf(acc[rt][1][r], acc[rt][1][r], q1);
      }
    // lanes 0-31 hold rows of column tile 0's columns, lanes 32-63 the other
    // rows: after the swap lane l < 32 has both halves of tile 0's column l
    // and lane l >= 32 both halves of tile 1's column l - 32 (column l of
    // the wave)
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(q0), __float_as_uint(q1),
                                                     false, false);
    const float qt = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);  // Q~ of this lane's column
    const float E = fmaf(scal[0], rm, fmaf(scal[1], am, scal[2]));
    // v_sqrt_f32 (within 1 ulp; the 4e-6 factors below cover it)
    const float sqt = __builtin_amdgcn_sqrtf(qt);
    // the slack and the 1 -+ 1e-5 factors (the rounding of the squares and
    // of the products) are in the kernel's constants
    const float rl = fmaxf(fmaf(sqt, 1.f - 4e-6f, -E), 0.f);
    const float rh = fmaf(sqt, 1.f + 4e-6f, E);
    hi = fmaf(scal[4], rl * rl, scal[3]);
    lo = fmaf(scal[6], rh * rh, scal[5]);
    if (!valid || n == diag_n) {
      hi = neg_inf();
      lo = neg_inf();
    }
  };

  // One pass over the block's kernels: screen, queue survivors from kernel
  // `resume` on (at most kScreenQ per lane; the first kernel not queued is
  // returned in next_resume, nb when all were), then compute the queued
  // pairs exactly.  keep_pass: survivors are the pairs that can pass the
  // keep test against the final normaliser cm.  Returns the lane's first
  // entry in its wave's list | (entries << 16).
  auto pass = [&](bool keep_pass, int64_t resume, int64_t &next_resume) -> int {
    int cnt = 0;
    next_resume = ke_;
    auto consider = [&](int64_t n, float hi, float lo) {
      bool surv;
      if (keep_pass) {
        surv = hi > cm + p.log_t_lo;
      } else {
        B = fmaxf(B, lo);
        // NaN bounds survive (and propagate through the exact values)
        surv = valid && n != diag_n && !(hi <= B - sp.margin);
      }
      if (surv && n >= resume && next_resume == ke_) {
        if (cnt < kScreenQ) {
          sq[cnt][tid] = (int16_t)n;
          ++cnt;
        } else {
          next_resume = n;
        }
      }
    };
    auto stage_kernels = [&](int64_t n0) -> int {
      return n0 >= ke_ ? 0 : (ke_ - n0 < kScreenKps ? (int)(ke_ - n0) : kScreenKps);
    };
    screen_stage_copy<D>(sp.frag + kb_ * FB, ring0, stage_kernels(kb_), w, lane);
    __syncthreads();
    // one stage: copy the next into the other array, screen this one's kernels
    auto stage = [&](int64_t n0, const uint8_t *fb, uint8_t *next) {
      const int64_t nx = n0 + kScreenKps;
      screen_stage_copy<D>(sp.frag + nx * FB, next, stage_kernels(nx), w, lane);
#pragma unroll
      for (int kk = 0; kk < kScreenKps; ++kk) {
        if (n0 + kk < ke_) {
          float hi, lo;
          bounds(fb + kk * L::kOne, n0 + kk, hi, lo);
          consider(n0 + kk, hi, lo);
        }
      }
      __syncthreads();
    };
    for (int64_t n0 = kb_; n0 < ke_; n0 += 2 * kScreenKps) {
      stage(n0, ring0, ring1);
      if (n0 + kScreenKps < ke_) stage(n0 + kScreenKps, ring1, ring0);
    }
    if constexpr (SPLIT) {
      // the usual case: no queue of the workgroup filled up, so the range's
      // survivors are all queued; each is appended to its kernel's list,
      // evaluated there by estep_screen_exact_kernel with the kernel's
      // record read once per wave, and combined per column later
      if (!keep_pass && resume == 0 && !__syncthreads_or(next_resume < ke_)) {
        if (valid) {
          const int64_t o = (int64_t)blockIdx.y * p.N + m;
          sp.part_m[o] = neg_inf();
          sp.part_s[o] = 0.f;
          sp.kcount[o] = 0;
          sp.sv_cnt[o] = (uint8_t)cnt;
        }
        // list slots: within the workgroup from per-kernel LDS counters,
        // then one global atomic per kernel of the range reserves them (the
        // stage arrays, free now, hold the counters, bases and local slots
        // when they are large enough)
        const int kb = (int)kb_, len = (int)(ke_ - kb_);
        int32_t *lcnt = reinterpret_cast<int32_t *>(kAggRing ? ring0 : agg_own);
        int32_t *gbase = lcnt + kScreenMaxRange;
        int16_t(*lslot)[256] =
            reinterpret_cast<int16_t(*)[256]>(kAggRing ? ring1 : agg_own + kAggCnt);
        auto put = [&](int k, int32_t n, int32_t slot) {
          const bool in = slot < sp.kl_cap;
          if (in) sp.kl_m[(int64_t)n * sp.kl_cap + slot] = (int32_t)m;
          sp.sv_e[(m * sp.nsplit + blockIdx.y) * kScreenQ + k] = ((in ? slot + 1 : 0) << 15) | n;
        };
        if (sc_agg_slots(D) && len <= kScreenMaxRange) {
          for (int t = tid; t < len; t += 256) lcnt[t] = 0;
          __syncthreads();
          for (int k = 0; k < cnt; ++k) lslot[k][tid] = (int16_t)atomicAdd(&lcnt[sq[k][tid] - kb], 1);
          __syncthreads();
          for (int t = tid; t < len; t += 256)
            if (lcnt[t] > 0) gbase[t] = atomicAdd(&sp.kl_cnt[kb + t], lcnt[t]);
          __syncthreads();
          for (int k = 0; k < cnt; ++k) {
            const int32_t n = sq[k][tid];
            put(k, n, gbase[n - kb] + lslot[k][tid]);
          }
        } else {
          // four list slots claimed at a time (their atomics in flight together)
          for (int k0 = 0; k0 < cnt; k0 += 4) {
            int32_t slot[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (k0 + i < cnt) slot[i] = atomicAdd(&sp.kl_cnt[sq[k0 + i][tid]], 1);
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (k0 + i < cnt) put(k0 + i, sq[k0 + i][tid], slot[i]);
          }
        }
        return -1;
      }
    }
    // exact phase (wave-local): compact the lanes' queues into one list,
    // one pair per lane, in kernel order per lane
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(incl, d, kWave);
      if (lane >= d) incl += t;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, kWave - 1, kWave);
    ScreenWaveList &W = wl[w];
    for (int k = 0; k < cnt; ++k) W.ent[excl + k] = ((int32_t)sq[k][tid] << 8) | lane;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int j0 = 0; j0 < total; j0 += kWave) {
      const int j = j0 + lane;
      if (j < total) {
        const int32_t e = W.ent[j];
        const int64_t pn = e >> 8;
        const int64_t pm = base + (e & 63);
        W.ent[j] = __float_as_int(log_rho_lane<D>(p.rec + pn * R, p.X + pm * D));
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return excl | (cnt << 16);
  };

  const int64_t cseg = chunk / p.seg_len;
  auto keep = [&](int64_t n, float x) {
    const float d = __fsub_rn(x, cm);
    float r;
    if (keep_resp(d, p.threshold, p.log_t_lo, &r)) {
      atomicOr((unsigned long long *)&p.masks[chunk * nb + n], 1ull << lane);
      atomicAdd(&p.seg_tot[cseg * nb + n], 1);
      atomicAdd(&p.row_tot[n], 1);
    }
  };

  // log-sum-exp passes (ACCUM, UPDATE), then keep passes (FIXED, or an
  // UPDATE block with a full keep queue); one pass call site
  bool keep_phase = MODE == LAKDE_ESTEP_FIXED;
  int64_t resume = 0;
  while (true) {
    int64_t next = ke_;
    const int r = pass(keep_phase, resume, next);
    if (SPLIT && r < 0) return;  // deferred to the combine kernel
    const int excl = r & 0xffff, cnt = r >> 16;
    const ScreenWaveList &W = wl[w];
    if (keep_phase) {
      if (valid)
        for (int k = 0; k < cnt; ++k) keep(sq[k][tid], __int_as_float(W.ent[excl + k]));
    } else {
      for (int k = 0; k < cnt; ++k) lse.add1(__int_as_float(W.ent[excl + k]));
      if (MODE == LAKDE_ESTEP_UPDATE) {
        // keep candidates: exact values that can pass the keep test against
        // any normaliser >= max(B, the largest exact value)
        const float Bk = fmaxf(B, lse.m);
        for (int k = 0; k < cnt; ++k) {
          const float x = __int_as_float(W.ent[excl + k]);
          if (__fsub_rn(x, Bk) > p.log_t_lo) {
            if (kcnt < kKeepQ) {
              kq_n[kcnt][tid] = sq[k][tid];
              kq_x[kcnt][tid] = x;
            }
            ++kcnt;
          }
        }
      }
      // the largest exact value bounds the normaliser too (later passes)
      B = fmaxf(B, lse.m);
    }
    if (__syncthreads_or(next < ke_)) {
      resume = next;
      continue;
    }
    if (keep_phase || MODE == LAKDE_ESTEP_FIXED) return;
    if constexpr (SPLIT) {
      // this range's part: estep_screen_combine_kernel finishes the column
      if (valid) {
        const int64_t o = (int64_t)blockIdx.y * p.N + m;
        sp.part_m[o] = lse.m;
        sp.part_s[o] = lse.s;
        sp.kcount[o] = (uint8_t)(kcnt < 255 ? kcnt : 255);
        sp.sv_cnt[o] = 0;
        const int kk = kcnt < kKeepQ ? kcnt : kKeepQ;
        for (int k = 0; k < kk; ++k) {
          const int64_t q = ((int64_t)blockIdx.y * kKeepQ + k) * p.N + m;
          sp.kq_n[q] = kq_n[k][tid];
          sp.kq_x[q] = kq_x[k][tid];
        }
      }
      return;
    }
    // the log-sum-exp is complete: the normaliser update
    const float l = lse.value();
    if (valid) {
      cm = MODE == LAKDE_ESTEP_ACCUM ? logaddexp_f(cm, l) : logaddexp_f(lb, l);
      p.c[m] = cm;
    }
    if (MODE == LAKDE_ESTEP_ACCUM) return;
    if (MODE == LAKDE_ESTEP_UPDATE && p.lazy_unsafe != nullptr)
      lazy_unsafe_test(p, valid, m, valid ? p.c_old[m] : 0.f, cm);
    if (!__syncthreads_or(kcnt > kKeepQ)) {
      if (valid)
        for (int k = 0; k < kcnt; ++k) keep(kq_n[k][tid], kq_x[k][tid]);
      return;
    }
    keep_phase = true;
    resume = 0;
  }
}

// ----------------------------------------------------- predictive density
// log p(y_m) = -log N + LSE_n t_nm, t_nm = c_n - h_n log1p(Q_nm) (the
// predictive constants of the record, lakde/local_adaptive_kde.py:209-223),
// with the screen above: every (kernel, test point) pair is bounded on the
// BF16 matrix cores (the same fragments and error bound E; t is decreasing
// in Q, so Q in [(sqrt Q~ - E)^2, (sqrt Q~ + E)^2] bounds t), a pair whose
// upper bound lies more than `margin` below a running lower bound B of its
// split's partial log-sum-exp (the largest lower bound so far, then the
// largest exact term) is dropped (all dropped terms of a split together are
// below n_split e^-margin of its partial sum: the host sets that to 1e-6),
// and the others are queued per
// test point and evaluated exactly in FP32 (log_pred<D>, the terms of the
// FP32 predictive kernel).  A wave evaluates its queues whenever one of
// them could overflow in the next stage, and at the end of the split.
// Grid: (test-point groups of 256, kernel splits); partial (max, sum-exp)
// per split, combined by pred_combine_kernel.
constexpr int kPredQ = 12;  // queued pairs per test point between evaluations

struct PredWaveList {
  int32_t ent[kPredQ * kWave];
};

// the exact term of one (kernel, test point) pair; not inlined, so that its
// registers (the point, the record values in flight) are not live across
// the screening loop of the caller, where they would force spills
template <int D>
__device__ __attribute__((noinline)) float log_pred_lane(const float *rec,
                                                         const float *__restrict__ yrow) {
  const float *r = reinterpret_cast<const float *>(__builtin_assume_aligned(rec, 16));
  float yv[D];
#pragma unroll
  for (int t = 0; t < D; ++t) yv[t] = yrow[t];
  return log_pred<D>(r, yv);
}

template <int D>
__global__ void __launch_bounds__(256, 3)
pred_screen_kernel(PParams p, const uint8_t *__restrict__ frag, const __bf16 *__restrict__ yb,
                   const float *__restrict__ ynorm, float margin) {
  using L = ScreenLds<D>;
  constexpr int Kp = sc_kp(D);
  constexpr int RT = sc_dp(D) / 32, KS = Kp / 16;
  constexpr int64_t R = rec_floats(D);
  constexpr int64_t FB = sc_frag_bytes(D);
  // two stage arrays (see estep_screen_kernel); the queues and the waves'
  // lists are live across stages, so they have their own memory
  __shared__ __attribute__((aligned(1024))) uint8_t ring0[L::kBytes];
  __shared__ __attribute__((aligned(1024))) uint8_t ring1[L::kBytes];
  __shared__ int32_t sq[kPredQ][256];
  __shared__ PredWaveList wl[4];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int64_t base = (int64_t)blockIdx.x * 256 + w * kWave;
  const int64_t m = base + lane;
  const bool valid = m < p.M;
  const int64_t split = blockIdx.y;
  const int64_t sbeg = split * p.per_split;
  const int64_t send = (sbeg + p.per_split) < p.N ? (sbeg + p.per_split) : p.N;
  const int64_t nk = send - sbeg;
  const uint8_t *fsplit = frag + sbeg * FB;

  // the lane's B fragments; reloaded after every exact evaluation (which
  // needs the registers) instead of being held through it
  sc_bf16x8 xf[2][KS];
  auto load_xf = [&]() {
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int64_t col = base + ct * 32 + (lane & 31);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        if (col < p.M)
          xf[ct][s] =
              *reinterpret_cast<const sc_bf16x8 *>(yb + col * Kp + s * 16 + 8 * (lane >> 5));
        else
          xf[ct][s] = __builtin_bit_cast(sc_bf16x8, (sc_u32x4){0u, 0u, 0u, 0u});
      }
    }
  };
  load_xf();
  const float am = valid ? ynorm[m] : 0.f;
  const float rm = valid ? ynorm[p.M + m] : 0.f;
  float B = neg_inf();
  Lse lse;
  lse.init();
  int cnt = 0;

  // bounds [lo, hi] of the exact term of kernel n (LDS image fb) at this
  // lane's test point
  auto bounds = [&](const uint8_t *fb, int64_t n, float &hi, float &lo) {
    const float *scal = reinterpret_cast<const float *>(fb + sc_ntiles(D) * 1024);
    const float *rc = p.rec + n * R + rec_c(D);
    const float cst = rc[2], h = rc[3];
    sc_f32x16 acc[RT][2];
    sc_init_acc<D>(acc, scal + kScreenScal, lane);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        if (!sc_tile_used(rt, s)) continue;
        const sc_bf16x8 a = *reinterpret_cast<const sc_bf16x8 *>(
            fb + sc_tile_index(D, rt, s) * 1024 + lane * 16);
        acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xf[0][s], acc[rt][0], 0, 0, 0);
        acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xf[1][s], acc[rt][1], 0, 0, 0);
      }
    }
    float q0 = 0.f, q1 = 0.f;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        q0 = fmaf(acc[rt][0][r], acc[rt][0][r], q0);
        q1 = fmaf(acc[rt][1][r], acc[rt][1][r], q1);
      }
    // the halves of each test point's sum of squares (see estep_screen_kernel)
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(q0), __float_as_uint(q1),
                                                     false, false);
    const float qt = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
    const float E = fmaf(scal[0], rm, fmaf(scal[1], am, scal[2]));
    const float sqt = __builtin_amdgcn_sqrtf(qt);  // within 1 ulp, covered below
    const float rl = fmaf(sqt, 1.f - 4e-6f, -E);
    const float qlb = rl > 0.f ? rl * rl * (1.f - 1e-5f) : 0.f;
    const float rh = fmaf(sqt, 1.f + 4e-6f, E);
    const float qub = rh * rh * (1.f + 1e-5f);
    // log1p(q) as ln 2 * v_log_f32(1 + q): the rounding of 1 + q (2^-24)
    // and the 1-ulp logarithm move h log1p by far less than the slack, which
    // also covers the rounding of the products and of the exact term
    constexpr float kLn2 = 0.693147182f;
    const float lhi = h * (kLn2 * __builtin_amdgcn_logf(1.f + qub));
    const float llo = h * (kLn2 * __builtin_amdgcn_logf(1.f + qlb));
    const float slack = fmaf(1e-4f, fabsf(cst) + lhi, 0.01f);
    hi = (cst - llo) + slack;
    lo = (cst - lhi) - slack;
    if (!valid) {
      hi = neg_inf();
      lo = neg_inf();
    }
  };

  // exact evaluation of the wave's queued pairs (wave-local), one per lane,
  // added to the owning lanes' log-sum-exps
  auto evaluate = [&]() {
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(incl, d, kWave);
      if (lane >= d) incl += t;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, kWave - 1, kWave);
    PredWaveList &W = wl[w];
    // entry: (queue slot << 6) | owning lane; replaced by the pair's term
    for (int k = 0; k < cnt; ++k) W.ent[excl + k] = (k << 6) | lane;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int j0 = 0; j0 < total; j0 += kWave) {
      const int j = j0 + lane;
      if (j < total) {
        const int e = W.ent[j];
        const int owner = e & (kWave - 1);
        const int64_t n = sq[e >> 6][w * kWave + owner];
        W.ent[j] = __float_as_int(log_pred_lane<D>(p.rec + n * R, p.Y + (base + owner) * D));
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = 0; k < cnt; ++k) lse.add1(__int_as_float(W.ent[excl + k]));
    B = fmaxf(B, lse.m);
    cnt = 0;
    __builtin_amdgcn_wave_barrier();
  };

  auto stage_kernels = [&](int64_t k0) -> int {
    return k0 >= nk ? 0 : (nk - k0 < kScreenKps ? (int)(nk - k0) : kScreenKps);
  };
  screen_stage_copy<D>(fsplit, ring0, stage_kernels(0), w, lane);
  __syncthreads();
  auto stage = [&](int64_t k0, const uint8_t *fb, uint8_t *nxt) {
    const int64_t kx = k0 + kScreenKps;
    screen_stage_copy<D>(fsplit + kx * FB, nxt, stage_kernels(kx), w, lane);
#pragma unroll
    for (int kk = 0; kk < kScreenKps; ++kk) {
      if (k0 + kk < nk) {
        const int64_t n = sbeg + k0 + kk;
        float hi, lo;
        bounds(fb + kk * L::kOne, n, hi, lo);
        B = fmaxf(B, lo);
        // NaN bounds survive (and propagate through the exact terms)
        if (valid && !(hi <= B - margin)) {
          sq[cnt][tid] = (int32_t)n;
          ++cnt;
        }
      }
    }
    if (__builtin_amdgcn_ballot_w64(cnt > kPredQ - kScreenKps) != 0) {
      evaluate();
      load_xf();
    }
    __syncthreads();
  };
  for (int64_t k0 = 0; k0 < nk; k0 += 2 * kScreenKps) {
    stage(k0, ring0, ring1);
    if (k0 + kScreenKps < nk) stage(k0 + kScreenKps, ring1, ring0);
  }
  if (__builtin_amdgcn_ballot_w64(cnt > 0) != 0) evaluate();
  if (valid) {
    p.part_m[split * p.M + m] = lse.m;
    p.part_s[split * p.M + m] = lse.s;
  }
}

template <int D>
int launch_pred_screen(const PParams &p, const uint8_t *frag, const __bf16 *yb,
                       const float *ynorm, float margin, hipStream_t s) {
  if constexpr (D < 8) {
    return LAKDE_EINVAL;
  } else {
    const dim3 grid((unsigned)ceil_div(p.M, 256), (unsigned)p.nsplit);
    hipLaunchKernelGGL((pred_screen_kernel<D>), grid, dim3(256), 0, s, p, frag, yb, ynorm, margin);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

// The work items of the kernels' lists, dense: one workgroup scans the
// per-kernel counts (each thread a run of kernels, then a scan of the runs)
// and writes (kernel << 8 | chunk) for every 256 entries of every list.
template <int D>  // (D: a copy per translation unit)
__global__ void __launch_bounds__(1024) screen_items_kernel(int64_t nb, SParams sp) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nb + 1023) / 1024;
  const int64_t n0 = t * per, n1 = n0 + per < nb ? n0 + per : nb;
  auto chunks = [&](int64_t n) {
    const int c = sp.kl_cnt[n];
    return ((c < sp.kl_cap ? c : sp.kl_cap) + 255) / 256;
  };
  int own = 0;
  for (int64_t n = n0; n < n1; ++n) own += chunks(n);
  part[t] = own;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int o = part[t] - own;
  for (int64_t n = n0; n < n1; ++n) {
    const int c = chunks(n);
    for (int ch = 0; ch < c; ++ch) sp.items[o++] = (int32_t)(n << 8) | ch;
  }
  if (t == 1023) sp.items_total[0] = part[1023];
}

// The exact values of the pairs in the kernels' lists: the work items over
// a fixed grid of workgroups; a workgroup copies the item's kernel record
// into LDS once (all of it in flight together) and its lanes, the entries'
// columns, read the factor as broadcast LDS loads (log_rho<D> as
// everywhere: the same bits).
constexpr int kExactGrid = 1024;

template <int D>
__global__ void __launch_bounds__(256) estep_screen_exact_kernel(const float *__restrict__ recs,
                                                                 const float *__restrict__ X,
                                                                 SParams sp) {
  constexpr int64_t R = rec_floats(D);  // a multiple of 4
  __shared__ sc_f32x4 srec[R / 4];
  const int total = sp.items_total[0];
  for (int it = blockIdx.x; it < total; it += gridDim.x) {
    const int32_t e = sp.items[it];
    const int64_t n = e >> 8;
    const int j0 = (e & 255) * 256;
    const int c = sp.kl_cnt[n];
    const int cnt = c < sp.kl_cap ? c : sp.kl_cap;
    const sc_f32x4 *g = reinterpret_cast<const sc_f32x4 *>(recs + n * R);
    __syncthreads();  // the previous item's reads of srec are done
    for (int i = threadIdx.x; i < (int)(R / 4); i += 256) srec[i] = g[i];
    const int j = j0 + threadIdx.x;
    float xv[D];
    int64_t m = 0;
    if (j < cnt) {
      m = sp.kl_m[n * sp.kl_cap + j];
#pragma unroll
      for (int q = 0; q < D; ++q) xv[q] = X[m * D + q];
    }
    __syncthreads();
    if (j < cnt) sp.kl_x[n * sp.kl_cap + j] = log_rho<D>(reinterpret_cast<const float *>(srec), xv);
  }
}

template <int D>  // (D: a copy per translation unit)
__global__ void __launch_bounds__(256) screen_zero_i32_kernel(int32_t *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Completes a split UPDATE E-step for 64 columns (one wave): each column's
// survivors from all ranges in range, then queue order (kernel order) --
// their exact values from the kernels' lists, or, for pairs whose kernel
// list was full, evaluated here (all such pairs of the 64 columns in one
// list, one pair per lane) -- and the ranges' own partial log-sum-exps
// merged into the new normaliser (the same update as the unsplit kernel),
// the lazy rebalance's unsafe-step test, and the keep tests of the exact
// values against the final normaliser.  A column with more keep
// candidates in a finished range than that range kept, or a wave with more
// pairs to evaluate than its list holds, is marked in `only` and finished
// by a FIXED keep pass over all kernels instead.
constexpr int kCombineList = 1024;  // pairs per wave list (4 KB of LDS)
constexpr int kScreenListCap = 16384;  // pairs per kernel list (a multiple of 256, at most 65536)

constexpr int kMaxSplit = 8;  // ranges of a split E-step at most

template <int D>
__global__ void __launch_bounds__(64) estep_screen_combine_kernel(EParams p, SParams sp) {
  constexpr int64_t R = rec_floats(D);
  __shared__ int32_t ent[kCombineList];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kWave;
  const int64_t m = base + lane;
  const bool valid = m < p.N;
  const int S = sp.nsplit;
  auto svi = [&](int r, int k) -> int64_t { return (m * S + r) * kScreenQ + k; };
  // the column's counts, partial sums and first entries of every range,
  // loaded together
  int cr[kMaxSplit], kq[kMaxSplit];
  float pm[kMaxSplit], ps[kMaxSplit];
#pragma unroll
  for (int r = 0; r < kMaxSplit; ++r) {
    const bool in = valid && r < S;
    cr[r] = in ? sp.sv_cnt[r * p.N + m] : 0;
    kq[r] = in ? sp.kcount[r * p.N + m] : 0;
    pm[r] = in ? sp.part_m[r * p.N + m] : neg_inf();
    ps[r] = in ? sp.part_s[r * p.N + m] : 0.f;
  }
  const float co = valid ? p.c_old[m] : 0.f;
  const float lb = valid ? logsubexp_f(co, logf(p.old_colsum[m]) + co) : 0.f;
  constexpr int KR = 2;  // entries per range held in registers
  int32_t e4[kMaxSplit][KR];
#pragma unroll
  for (int r = 0; r < kMaxSplit; ++r)
#pragma unroll
    for (int k = 0; k < KR; ++k) e4[r][k] = k < cr[r] ? sp.sv_e[svi(r, k)] : 0;
  Lse t;
  t.init();
  bool over = false;
  int nf = 0;  // this column's pairs to evaluate here (their kernel list was full)
#pragma unroll
  for (int r = 0; r < kMaxSplit; ++r) {
    t.merge(pm[r], ps[r]);
    over = over || kq[r] > kKeepQ;
#pragma unroll
    for (int k = 0; k < KR; ++k) nf += (k < cr[r] && (e4[r][k] >> 15) == 0) ? 1 : 0;
    for (int k = KR; k < cr[r]; ++k) nf += (sp.sv_e[svi(r, k)] >> 15) == 0;
  }
  auto listed = [&](int32_t e) -> float {
    return sp.kl_x[(int64_t)(e & 0x7fff) * sp.kl_cap + ((e >> 15) - 1)];
  };
  const int64_t chunk = m / kWave;
  const int64_t cseg = chunk / p.seg_len;
  auto keep = [&](int64_t n, float x, float cm) {
    const float d = __fsub_rn(x, cm);
    float rr;
    if (keep_resp(d, p.threshold, p.log_t_lo, &rr)) {
      atomicOr((unsigned long long *)&p.masks[chunk * p.nb + n], 1ull << lane);
      atomicAdd(&p.seg_tot[cseg * p.nb + n], 1);
      atomicAdd(&p.row_tot[n], 1);
    }
  };
  auto keep_queued = [&](int r, float cm) {
    for (int k = 0; k < kq[r]; ++k) {
      const int64_t q = ((int64_t)r * kKeepQ + k) * p.N + m;
      keep(sp.kq_n[q], sp.kq_x[q], cm);
    }
  };

  if (__ballot(nf > 0) == 0) {
    // every value is in the kernels' lists: in range, then queue order
#pragma unroll
    for (int r = 0; r < kMaxSplit; ++r) {
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (k < cr[r]) t.add1(listed(e4[r][k]));
      for (int k = KR; k < cr[r]; ++k) t.add1(listed(sp.sv_e[svi(r, k)]));
    }
    float cm = 0.f;
    if (valid) {
      cm = logaddexp_f(lb, t.value());
      p.c[m] = cm;
      sp.only[m] = over ? 1 : 0;
    }
    if (p.lazy_unsafe != nullptr) lazy_unsafe_test(p, valid, m, co, cm);
    if (!valid || over) return;
#pragma unroll
    for (int r = 0; r < kMaxSplit; ++r) {
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (k < cr[r]) keep(e4[r][k] & 0x7fff, listed(e4[r][k]), cm);
      for (int k = KR; k < cr[r]; ++k) {
        const int32_t e = sp.sv_e[svi(r, k)];
        keep(e & 0x7fff, listed(e), cm);
      }
      if (r < S) keep_queued(r, cm);
    }
    return;
  }

  // some pairs of the wave are evaluated here, one per lane, in windows of
  // kCombineList (one window but in dense regions)
  int incl = nf;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int u = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += u;
  }
  const int excl = incl - nf;
  const int total = __shfl(incl, kWave - 1, kWave);
  // value of entry e: from its kernel's list, or the next of this column's
  // evaluated pairs (list position f)
  auto value = [&](int32_t e, int &f, int w0) -> float {
    if ((e >> 15) != 0) return listed(e);
    return __int_as_float(ent[(f++) - w0]);
  };
  // the column's entries in order, up to the first evaluated pair at list
  // position >= lim
  int rc = 0, kc = 0, fc = excl;
  auto walk = [&](int lim, int w0) {
    while (rc < S) {
      if (kc >= sp.sv_cnt[rc * p.N + m]) {
        ++rc;
        kc = 0;
        continue;
      }
      const int32_t e = sp.sv_e[svi(rc, kc)];
      if ((e >> 15) == 0 && fc >= lim) return;
      t.add1(value(e, fc, w0));
      ++kc;
    }
  };
  for (int w0 = 0; w0 < total; w0 += kCombineList) {
    if (valid && nf > 0 && excl < w0 + kCombineList && excl + nf > w0) {
      int o = excl;
      for (int r = 0; r < S; ++r) {
        const int c = sp.sv_cnt[r * p.N + m];
        for (int k = 0; k < c; ++k) {
          const int32_t e = sp.sv_e[svi(r, k)];
          if ((e >> 15) != 0) continue;
          if (o >= w0 && o < w0 + kCombineList) ent[o - w0] = ((e & 0x7fff) << 6) | lane;
          ++o;
        }
      }
    }
    __syncthreads();
    const int wn = total - w0 < kCombineList ? total - w0 : kCombineList;
    for (int j = lane; j < wn; j += kWave) {
      const int32_t e = ent[j];
      ent[j] = __float_as_int(log_rho_lane<D>(p.rec + (int64_t)(e >> 6) * R, p.X + (base + (e & 63)) * D));
    }
    __syncthreads();
    if (valid) walk(w0 + kCombineList, w0);
    __syncthreads();
  }
  if (valid) walk(0x7fffffff, 0);
  // the evaluated values stay for the keep tests only when one window held
  // them all
  over = over || total > kCombineList;
  float cm = 0.f;
  if (valid) {
    cm = logaddexp_f(lb, t.value());
    p.c[m] = cm;
    sp.only[m] = over ? 1 : 0;
  }
  if (p.lazy_unsafe != nullptr) lazy_unsafe_test(p, valid, m, co, cm);
  if (!valid || over) return;
  int f = excl;
  for (int r = 0; r < S; ++r) {
    const int c = sp.sv_cnt[r * p.N + m];
    for (int k = 0; k < c; ++k) {
      const int32_t e = sp.sv_e[svi(r, k)];
      keep(e & 0x7fff, value(e, f, 0), cm);
    }
    keep_queued(r, cm);
  }
}

// The split UPDATE form (sp.nsplit > 1; the block's kernels in nsplit
// ranges, one workgroup per (column group, range)): the ranges' screens,
// the per-column combine, and a FIXED keep pass over all kernels for the
// columns whose keep candidates overflowed (its workgroups without such
// columns return at once).  A launch of ceil(N / 256) workgroups of the
// unsplit kernel is about five waves of workgroups at three per CU, and the
// last, partial wave costs as much as a full one; with the kernels split in
// ranges the launch is many more, shorter workgroups.
template <int D>
int launch_estep_screen(const EParams &p, const SParams &sp, int mode, hipStream_t s) {
  if constexpr (D < 8) {
    return LAKDE_EINVAL;
  } else {
    const dim3 grid((unsigned)ceil_div(p.N, 256));
    if (mode == LAKDE_ESTEP_UPDATE && sp.nsplit > 1) {
      hipLaunchKernelGGL((screen_zero_i32_kernel<D>), dim3((unsigned)ceil_div(p.nb, 256)), dim3(256), 0,
                         s, sp.kl_cnt, p.nb);
      hipLaunchKernelGGL((estep_screen_kernel<D, LAKDE_ESTEP_UPDATE, true>),
                         dim3(grid.x, (unsigned)sp.nsplit), dim3(256), 0, s, p, sp);
      hipLaunchKernelGGL((screen_items_kernel<D>), dim3(1), dim3(1024), 0, s, p.nb, sp);
      hipLaunchKernelGGL((estep_screen_exact_kernel<D>), dim3(kExactGrid), dim3(256), 0, s, p.rec,
                         p.X, sp);
      hipLaunchKernelGGL((estep_screen_combine_kernel<D>), dim3((unsigned)ceil_div(p.N, kWave)),
                         dim3(kWave), 0, s, p, sp);
      hipLaunchKernelGGL((estep_screen_kernel<D, LAKDE_ESTEP_FIXED>), grid, dim3(256), 0, s, p,
                         sp);
      return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
    }
    if (mode == LAKDE_ESTEP_ACCUM)
      hipLaunchKernelGGL((estep_screen_kernel<D, LAKDE_ESTEP_ACCUM>), grid, dim3(256), 0, s, p, sp);
    else if (mode == LAKDE_ESTEP_FIXED)
      hipLaunchKernelGGL((estep_screen_kernel<D, LAKDE_ESTEP_FIXED>), grid, dim3(256), 0, s, p, sp);
    else
      hipLaunchKernelGGL((estep_screen_kernel<D, LAKDE_ESTEP_UPDATE>), grid, dim3(256), 0, s, p, sp);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

template <int D>
int launch_pack_screen(const float *rec, int64_t nb, const float *origin, uint8_t *frag,
                       hipStream_t s) {
  if constexpr (D < 8) {
    return LAKDE_EINVAL;
  } else {
    if (nb == 0) return LAKDE_OK;
    hipLaunchKernelGGL((pack_screen_kernel<D>), dim3((unsigned)nb), dim3(64), 0, s, rec, origin,
                       frag);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

template <int D>
int launch_screen_points(const float *X, int64_t N, const float *origin, __bf16 *xb,
                         float *xnorm, hipStream_t s) {
  if constexpr (D < 8) {
    return LAKDE_EINVAL;
  } else {
    if (N == 0) return LAKDE_OK;
    hipLaunchKernelGGL((screen_points_kernel<D>), dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s,
                       X, N, origin, xb, xnorm);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

template <int D>
int launch_screen_knn_bound(const float *rec, const float *X, int64_t N, const void *graph,
                            int gdtype, int64_t gk, int J, float *out, hipStream_t s) {
  if constexpr (D < 8) {
    return LAKDE_EINVAL;
  } else {
    if (N == 0) return LAKDE_OK;
    hipLaunchKernelGGL((screen_knn_bound_kernel<D>), dim3((unsigned)ceil_div(N * J, 256)),
                       dim3(256), 0, s, rec, X, N, graph, gdtype, gk, J, out);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

#define LAKDE_DECL_SCREEN_D(d)                                                                \
  extern template int launch_screen_knn_bound<d>(const float *, const float *, int64_t,       \
                                                 const void *, int, int64_t, int, float *,    \
                                                 hipStream_t);                                 \
  extern template int launch_estep_screen<d>(const EParams &, const SParams &, int, hipStream_t); \
  extern template int launch_pack_screen<d>(const float *, int64_t, const float *, uint8_t *,  \
                                            hipStream_t);                                       \
  extern template int launch_screen_points<d>(const float *, int64_t, const float *, __bf16 *,  \
                                              float *, hipStream_t);                             \
  extern template int launch_pred_screen<d>(const PParams &, const uint8_t *, const __bf16 *,   \
                                            const float *, float, hipStream_t);

}  // namespace lakde
