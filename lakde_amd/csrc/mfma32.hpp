// Plain FP32 E-step on the matrix cores (17 <= D <= 63).
//
// The quadratic forms y = U_n (x_m - x_n) of every (kernel n, point m) pair
// of a block run on v_mfma_f32_32x32x2_f32 (FP32 products, FP32
// accumulation) in the difference form: the B operand x_m - x_n is formed
// per K-step on the vector ALUs from the point's coordinates (held in
// registers) and the kernel's centre.  Q = |y|^2 and log rho = const +
// expo Q are FP32 values for every pair, and the streaming log-sum-exp and
// the normaliser update take all of them, as the FP32 vector E-step does
// (the two differ by FP32 rounding order only).
//
// The keep test needs the final normaliser and the value the compaction
// recomputes (log_rho<D>, which sets the stored bits).  Both the matrix-core
// y~ and the vector form's y are within (K + 2) u sum_j |U_ij| |d_j| of the
// exact y per row (u = 2^-24, d = x_m - x_n rounded), so by Cauchy-Schwarz
// and |d| <= |x_m - o| + |x_n - o| (o a fixed origin)
//     | y~ - y_vector | <= E_nm = kf F_n (a_m + b_n),
// kf = 2 (2 (K + 2) u) >= the sum of both errors with a factor 2 to spare
// (F_n = |U_n|_F, a_m = |x_m - o|, b_n = |x_n - o|, all rounded up).  With
// the upper bound hi of log rho formed from sqrt(Q~) - E as in the screening
// E-step, a pair is a keep candidate when hi can pass the keep test against
// a lower bound B of the final normaliser (UPDATE: max of the old-mass
// bound and the largest value so far; FIXED: the given normaliser).
// Candidates (a few per column and block) are queued per lane in LDS,
// recomputed exactly with log_rho<D> one pair per lane, and tested against
// the final normaliser, so the masks carry exactly the compaction's
// decisions.  A lane whose queue fills flushes it (FIXED) or sends the
// block to a keep pass (UPDATE, where the normaliser is not final yet).
//
// MFMA layout (v_mfma_f32_32x32x2_f32): lane l holds A[l % 32][l / 32] and
// B[l / 32][l % 32]; C holds 16 floats per lane, column l % 32, row
// 8 (r / 4) + 4 (l / 32) + r % 4 for register r.  K-step s covers k = 2s and
// 2s + 1, so lane l reads k = 2s + l / 32: the fragment stores each row's
// even and odd columns in two halves (k -> (k & 1) KH + k / 2).
#pragma once

#include "screen.hpp"

namespace lakde {

typedef float m32_f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int m32_dp(int D) { return (D + 31) / 32 * 32; }
__host__ __device__ constexpr int m32_kh(int D) { return m32_dp(D) / 2; }
__host__ __device__ constexpr int m32_ld(int D) { return m32_dp(D) + 4; }
// fragment block per kernel (its LDS image): U [Dp rows x ld] with each
// row's columns split by parity, x_n [Dp] split the same way, scal[4] =
// const, expo, kf F_n, kf F_n b_n
__host__ __device__ constexpr int64_t m32_frag_floats(int D) {
  return (int64_t)m32_dp(D) * m32_ld(D) + m32_dp(D) + 4;
}
constexpr int kM32Q = 8;  // keep-candidate queue entries per lane

template <int D>
__global__ void __launch_bounds__(64) pack_m32_kernel(const float *__restrict__ rec,
                                                      const float *__restrict__ origin,
                                                      float *__restrict__ frag) {
  constexpr int Dp = m32_dp(D), KH = m32_kh(D), LD = m32_ld(D);
  constexpr int64_t R = rec_floats(D);
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x;
  const float *r = rec + n * R;
  float *f = frag + n * m32_frag_floats(D);
  for (int t = lane; t < Dp * LD; t += 64) {
    const int i = t / LD, c = t - i * LD;
    float v = 0.f;
    if (c < Dp) {
      const int j = c < KH ? 2 * c : 2 * (c - KH) + 1;
      if (i < D && j <= i) v = r[i * (i + 1) / 2 + j];
    }
    f[t] = v;
  }
  float *xn = f + Dp * LD;
  if (lane < Dp) {
    const int j = lane < KH ? 2 * lane : 2 * (lane - KH) + 1;
    xn[lane] = j < D ? r[rec_x(D) + j] : 0.f;
  }
  double f2 = 0.0, b2 = 0.0;
  if (lane < D) {
    const float *row = r + lane * (lane + 1) / 2;
    for (int j = 0; j <= lane; ++j) f2 += (double)row[j] * (double)row[j];
    const double dl = (double)r[rec_x(D) + lane] - (double)origin[lane];
    b2 = dl * dl;
  }
  f2 = wave_sum(f2);
  b2 = wave_sum(b2);
  if (lane == 0) {
    constexpr double kf = 4.0 * (D + 2) * 0x1p-24;
    const double F = sqrt(f2) * (1.0 + 1e-6), b = sqrt(b2) * (1.0 + 1e-6);
    float *scal = xn + Dp;
    scal[0] = r[rec_c(D) + 0];
    scal[1] = r[rec_c(D) + 1];
    scal[2] = (float)(kf * F * (1.0 + 1e-6));
    scal[3] = (float)(kf * F * b * (1.0 + 1e-6));
  }
}

// Cooperative copy of one fragment block into LDS: every load issued into
// registers before the current kernel's work, the stores after it.
template <int D>
struct M32Stage {
  static constexpr int kGran = (int)(m32_frag_floats(D) / 4);  // 16-byte granules
  static constexpr int kPer = (kGran + 255) / 256;
  sc_u32x4 g[kPer];
  __device__ __forceinline__ void load(const float *__restrict__ src) {
    const sc_u32x4 *s = reinterpret_cast<const sc_u32x4 *>(src);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int q = (int)threadIdx.x + 256 * i;
      if (q < kGran) g[i] = s[q];
    }
  }
  __device__ __forceinline__ void store(float *dst) const {
    sc_u32x4 *d = reinterpret_cast<sc_u32x4 *>(dst);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int q = (int)threadIdx.x + 256 * i;
      if (q < kGran) d[q] = g[i];
    }
  }
};

template <int D, int MODE>
__global__ void __launch_bounds__(256, 2)
estep_m32_kernel(EParams p, const float *__restrict__ frag, const float *__restrict__ xnorm) {
  constexpr int Dp = m32_dp(D), KH = m32_kh(D), LD = m32_ld(D), RT = Dp / 32;
  constexpr int KS = (D + 1) / 2;  // K-steps that reach a column < D
  constexpr int64_t R = rec_floats(D);
  constexpr int64_t FF = m32_frag_floats(D);
  __shared__ __attribute__((aligned(16))) float lds[2][FF];
  __shared__ int16_t kq[kM32Q][256];
  __shared__ ScreenWaveList wl[4];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int h = lane >> 5;  // the K half of this lane (k = 2s + h)
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int64_t base = (int64_t)blockIdx.x * 256 + w * kWave;  // the wave's 64 columns = one chunk
  const int64_t m = base + lane;
  const bool valid = m < p.N;
  const int64_t chunk = base / kWave;
  const int64_t nb = p.nb;
  const int64_t diag_n = m - p.row0;

  // this lane's B coordinates: column ct * 32 + lane % 32 of the wave, k = 2s + h
  float xmr[2][KS];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    const int64_t col = base + ct * 32 + (lane & 31);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 2 * s + h;
      xmr[ct][s] = (col < p.N && k < D) ? p.X[col * D + k] : 0.f;
    }
  }
  const float am = valid ? xnorm[m] : 0.f;

  float lb = neg_inf();  // UPDATE: bound of the normaliser from the old mass
  float cm = 0.f;        // ACCUM, FIXED: the given normaliser; later the new one
  if (MODE == LAKDE_ESTEP_UPDATE && valid) {
    const float co = p.c_old[m];
    lb = logsubexp_f(co, logf(p.old_colsum[m]) + co);
  }
  if (MODE != LAKDE_ESTEP_UPDATE && valid) cm = p.c[m];
  Lse lse;
  lse.init();
  int cnt = 0;  // queued keep candidates (UPDATE: may exceed kM32Q)

  // log rho of kernel n (LDS image U) at this lane's column, and the upper
  // bound hi of the exact value (see above); -inf for the diagonal and
  // invalid columns
  auto eval = [&](const float *U, int64_t n, float &x, float &hi) {
    const float *xn = U + Dp * LD;
    const float *scal = xn + Dp;
    sc_f32x16 acc[RT][2];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[rt][0][r] = 0.f;
        acc[rt][1][r] = 0.f;
      }
#pragma unroll
    for (int s4 = 0; s4 < (KS + 3) / 4; ++s4) {
      const m32_f32x4 xv = *reinterpret_cast<const m32_f32x4 *>(xn + h * KH + 4 * s4);
      m32_f32x4 av[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
        if (8 * s4 <= rt * 32 + 31)
          av[rt] = *reinterpret_cast<const m32_f32x4 *>(U + (rt * 32 + (lane & 31)) * LD +
                                                        h * KH + 4 * s4);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = 4 * s4 + i;
        if (s >= KS) continue;
        const float b0 = xmr[0][s] - xv[i], b1 = xmr[1][s] - xv[i];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          if (2 * s > rt * 32 + 31) continue;  // U is lower triangular
          acc[rt][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[rt][i], b0, acc[rt][0], 0, 0, 0);
          acc[rt][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[rt][i], b1, acc[rt][1], 0, 0, 0);
        }
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
    const bool upper = lane >= 32;
    const float recv = __shfl_xor(upper ? q0 : q1, 32, kWave);
    const float qt = (upper ? q1 : q0) + recv;  // Q~ of this lane's column
    const float cst = scal[0], expo = scal[1];
    x = __fadd_rn(__fmul_rn(qt, expo), cst);
    const float E = fmaf(scal[2], am, scal[3]);
    const float rl = fmaf(sqrtf(qt), 1.f - 4e-6f, -E);
    const float qlb = rl > 0.f ? rl * rl * (1.f - 1e-5f) : 0.f;
    hi = fmaf(expo, qlb, cst) + fmaf(1e-4f, fabsf(cst), 0.01f);
    if (!valid || n == diag_n) {
      x = neg_inf();
      hi = neg_inf();
    }
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
  // the lanes' queued candidates recomputed exactly (one pair per lane
  // through the wave's list) and kept or not by their own lanes
  auto flush = [&](int q) {
    int incl = q;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(incl, d, kWave);
      if (lane >= d) incl += t;
    }
    const int excl = incl - q;
    const int total = __shfl(incl, kWave - 1, kWave);
    ScreenWaveList &W = wl[w];
    for (int k = 0; k < q; ++k) W.ent[excl + k] = ((int32_t)kq[k][tid] << 8) | lane;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int j0 = 0; j0 < total; j0 += kWave) {
      const int j = j0 + lane;
      if (j < total) {
        const int32_t e = W.ent[j];
        W.ent[j] = __float_as_int(log_rho_lane<D>(p.rec + (int64_t)(e >> 8) * R,
                                                  p.X + (base + (e & 63)) * D));
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid)
      for (int k = 0; k < q; ++k) keep(kq[k][tid], __int_as_float(W.ent[excl + k]));
    __builtin_amdgcn_wave_barrier();
  };

  // One pass over the block's kernels.  LSE pass (ACCUM, UPDATE): every
  // value into the log-sum-exp; UPDATE also queues its keep candidates
  // against B.  Keep pass (FIXED, or an UPDATE block whose queue
  // overflowed): candidates against the final normaliser cm, flushed
  // whenever some lane's queue is full.
  auto pass = [&](bool keep_pass) {
    M32Stage<D> st;
    st.load(frag);
    st.store(lds[0]);
    __syncthreads();
    int buf = 0;
    for (int64_t n = 0; n < nb; ++n) {
      if (n + 1 < nb) st.load(frag + (n + 1) * FF);
      float x, hi;
      eval(lds[buf], n, x, hi);
      if (keep_pass) {
        if (__fsub_rn(hi, cm) > p.log_t_lo) {
          kq[cnt][tid] = (int16_t)n;
          ++cnt;
        }
        if (__ballot(cnt == kM32Q) != 0ull) {
          flush(cnt);
          cnt = 0;
        }
      } else {
        lse.add1(x);
        if (MODE == LAKDE_ESTEP_UPDATE) {
          const float B = fmaxf(lb, lse.m);
          if (__fsub_rn(hi, B) > p.log_t_lo) {
            if (cnt < kM32Q) kq[cnt][tid] = (int16_t)n;
            ++cnt;
          }
        }
      }
      if (n + 1 < nb) st.store(lds[buf ^ 1]);
      __syncthreads();
      buf ^= 1;
    }
  };

  if (MODE == LAKDE_ESTEP_FIXED) {
    pass(true);
    flush(cnt);
    return;
  }
  pass(false);
  const float l = lse.value();
  if (valid) {
    cm = MODE == LAKDE_ESTEP_ACCUM ? logaddexp_f(cm, l) : logaddexp_f(lb, l);
    p.c[m] = cm;
  }
  if (MODE == LAKDE_ESTEP_ACCUM) return;
  if (p.lazy_unsafe != nullptr) lazy_unsafe_test(p, valid, m, valid ? p.c_old[m] : 0.f, cm);
  if (!__syncthreads_or(cnt > kM32Q)) {
    flush(cnt);
    return;
  }
  cnt = 0;
  pass(true);
  flush(cnt);
}

template <int D>
int launch_estep_m32(const EParams &p, const float *frag, const float *xnorm, int mode,
                     hipStream_t s) {
  if constexpr (D < 17) {
    return LAKDE_EINVAL;
  } else {
    const dim3 grid((unsigned)ceil_div(p.N, 256));
    if (mode == LAKDE_ESTEP_ACCUM)
      hipLaunchKernelGGL((estep_m32_kernel<D, LAKDE_ESTEP_ACCUM>), grid, dim3(256), 0, s, p, frag,
                         xnorm);
    else if (mode == LAKDE_ESTEP_FIXED)
      hipLaunchKernelGGL((estep_m32_kernel<D, LAKDE_ESTEP_FIXED>), grid, dim3(256), 0, s, p, frag,
                         xnorm);
    else
      hipLaunchKernelGGL((estep_m32_kernel<D, LAKDE_ESTEP_UPDATE>), grid, dim3(256), 0, s, p,
                         frag, xnorm);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

template <int D>
int launch_pack_m32(const float *rec, int64_t nb, const float *origin, float *frag,
                    hipStream_t s) {
  if constexpr (D < 17) {
    return LAKDE_EINVAL;
  } else {
    if (nb == 0) return LAKDE_OK;
    hipLaunchKernelGGL((pack_m32_kernel<D>), dim3((unsigned)nb), dim3(64), 0, s, rec, origin,
                       frag);
    return hipGetLastError() == hipSuccess ? LAKDE_OK : LAKDE_EHIP;
  }
}

#define LAKDE_DECL_M32_D(d)                                                                    \
  extern template int launch_estep_m32<d>(const EParams &, const float *, const float *, int,  \
                                          hipStream_t);                                         \
  extern template int launch_pack_m32<d>(const float *, int64_t, const float *, float *,       \
                                         hipStream_t);

}  // namespace lakde
