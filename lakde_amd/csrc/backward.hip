// Backward of the streaming predictive density (differentiation form).
//
// Forward (lakde/local_adaptive_kde.py:200-223):
//   out_m = -ln N + LSE_n lp_nm,  lp_nm = coeff_n - h_n log1p(Q_nm),
//   Q_nm = ||Linv_n (y_m - x_n)||^2, h_n = (nu_n + 1) / 2,
//   coeff_n = lgamma(h_n) - lgamma(h_n - D/2) + sum_i log Linv_n[i,i] - D/2 ln pi.
// With upstream gradient g_m and w_nm = exp(lp_nm - LSE_m) recomputed from
// the saved out_m (no N x M weight tensor is stored):
//   d/dy_m       = sum_n a_nm 2 Linv_n^T z_nm,     a_nm = -g_m w_nm h_n / (1 + Q_nm)
//   d/dLinv_n    = tril(2 Linv_n S_n) + diag(beta_n / Linv_n[i,i]),
//   d/dx_n       = -2 Linv_n^T Linv_n s1_n,
//   d/dnu_n      = (beta_n (psi(h_n) - psi(h_n - D/2)) - tau_n) / 2,
// where S_n = sum_m a_nm d d^T, s1_n = sum_m a_nm d (d = y_m - x_n),
// beta_n = sum_m g_m w_nm and tau_n = sum_m g_m w_nm log1p(Q_nm).  This file
// computes dY (with the D-templated bwd_y kernel) and S, s1, beta, tau; the
// D x D products that turn them into d/dLinv etc. are left to the caller.
#include "fused_impl.hpp"
#include "record.hpp"

namespace lakde {

LAKDE_FOR_EACH_D(LAKDE_DECL_D)

static int dispatch_bwd_y(int64_t D, const PParams &p, const float *lse, const float *g,
                          float *part, hipStream_t s) {
  switch (D) {
#define CASE(d) case d: return launch_bwd_y<d>(p, lse, g, part, s);
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  return LAKDE_EINVAL;
}

// records from explicit factors: Linv (N x D x D, lower), centres, nu
__global__ void __launch_bounds__(64) pack_records_kernel(const float *__restrict__ Linv,
                                                         const float *__restrict__ Xc,
                                                         const float *__restrict__ nu,
                                                         int D, float *__restrict__ rec) {
  const int64_t n = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t R = rec_floats(D);
  const float *L = Linv + n * (int64_t)D * D;
  float *r = rec + n * R;
  // 8 loads in flight per lane
#pragma unroll 8
  for (int t = lane; t < (int)rec_tri(D); t += 64) {
    int i, c;
    tri_index(t, &i, &c);
    r[t] = L[i * D + c];
  }
  if (lane < D) r[rec_x(D) + lane] = Xc[n * D + lane];
  double ls = lane < D ? log((double)L[lane * D + lane]) : 0.0;
  ls = wave_sum(ls);
  if (lane == 0) record_consts(r + rec_c(D), nu[n], (float)(2.0 * ls), D);
}

__global__ void __launch_bounds__(256) lse_from_out_kernel(const float *out, int64_t M, float logn,
                                                          float *lse) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m < M) lse[m] = out[m] + logn;
}

__global__ void __launch_bounds__(256) sum_splits_kernel(const float *part, int64_t nsplit,
                                                        int64_t MD, float *out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= MD) return;
  float s = 0.f;
  for (int64_t k = 0; k < nsplit; ++k) s += part[k * MD + i];
  out[i] = s;
}

// One workgroup per kernel n: chunks of 64 test points; 4 threads per point
// form z = Linv d (rows strided by 4) and Q; then every thread accumulates
// a 3 x 3 tile of the upper triangle of S (ceil(D/3) (ceil(D/3) + 1) / 2 <=
// 231 tiles: 6 LDS reads per point for 9 products) and, for j < D, s1.
// Deterministic order.

__global__ void __launch_bounds__(256) bwd_kernel_stats(const float *__restrict__ rec, int D,
                                                       const float *__restrict__ Y, int64_t M,
                                                       const float *__restrict__ lse,
                                                       const float *__restrict__ g,
                                                       float *__restrict__ S,
                                                       float *__restrict__ s1,
                                                       float *__restrict__ beta,
                                                       float *__restrict__ tau) {
  __shared__ float Ls[64][65];
  __shared__ float dl[64][65];  // padded: a wave reads 16 rows at one column
  __shared__ float al[64];
  __shared__ float red[2][256];
  const int64_t n = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t R = rec_floats(D);
  const float *r = rec + n * R;
  for (int t = tid; t < D * D; t += 256) {
    const int i = t / D, j = t - i * D;
    Ls[i][j] = j <= i ? r[i * (i + 1) / 2 + j] : 0.f;
  }
  const float *xn = r + rec_x(D);
  const float coeff = r[rec_c(D) + 2];
  const float h = r[rec_c(D) + 3];
  // this thread's tile: rows 3 ti .., columns 3 tj .., ti <= tj; reads are
  // clamped into [0, D) (entries outside the triangle are not written)
  const int nT = (D + 2) / 3;
  const bool tile_ok = tid < nT * (nT + 1) / 2;
  int ti = 0, tj = 0;
  if (tile_ok) {
    int r = tid;
    while (r >= nT - ti) { r -= nT - ti; ++ti; }
    tj = ti + r;
  }
  int ir[3], jc[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    ir[u] = 3 * ti + u < D ? 3 * ti + u : D - 1;
    jc[u] = 3 * tj + u < D ? 3 * tj + u : D - 1;
  }
  float acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.f;
  float s1acc = 0.f, bacc = 0.f, tacc = 0.f;
  const int pt = tid >> 2, qt = tid & 3;
  for (int64_t m0 = 0; m0 < M; m0 += 64) {
    __syncthreads();
    for (int t = tid; t < 64 * D; t += 256) {
      const int p = t / D, j = t - p * D;
      const int64_t m = m0 + p;
      dl[p][j] = m < M ? Y[m * D + j] - xn[j] : 0.f;
    }
    __syncthreads();
    float qpart = 0.f;
    for (int i = qt; i < D; i += 4) {
      float y = 0.f;
      for (int j = 0; j <= i; ++j) y = fmaf(Ls[i][j], dl[pt][j], y);
      qpart = fmaf(y, y, qpart);
    }
    qpart += __shfl_xor(qpart, 1, kWave);
    qpart += __shfl_xor(qpart, 2, kWave);
    if (qt == 0) {
      const int64_t m = m0 + pt;
      float a = 0.f;
      if (m < M) {
        const float l1 = log1pf(qpart);
        const float lp = __fsub_rn(coeff, __fmul_rn(h, l1));
        const float gw = g[m] * expf(lp - lse[m]);
        a = -gw * h / (1.f + qpart);
        bacc += gw;
        tacc += gw * l1;
      }
      al[pt] = a;
    }
    __syncthreads();
    for (int p = 0; p < 64; ++p) {
      const float a = al[p];
      if (tile_ok) {
        float wi[3], dj[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          wi[u] = a * dl[p][ir[u]];
          dj[u] = dl[p][jc[u]];
        }
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
          for (int b2 = 0; b2 < 3; ++b2) acc[a2 * 3 + b2] = fmaf(wi[a2], dj[b2], acc[a2 * 3 + b2]);
      }
      if (tid < D) s1acc = fmaf(a, dl[p][tid], s1acc);
    }
  }
  const int64_t DD = (int64_t)D * D;
  if (tile_ok) {
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
      for (int b2 = 0; b2 < 3; ++b2) {
        const int i = 3 * ti + a2, j = 3 * tj + b2;
        if (i < D && j < D && i <= j) {
          S[n * DD + i * D + j] = acc[a2 * 3 + b2];
          S[n * DD + j * D + i] = acc[a2 * 3 + b2];
        }
      }
  }
  if (tid < D) s1[n * D + tid] = s1acc;
  red[0][tid] = bacc;
  red[1][tid] = tacc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      red[0][tid] += red[0][tid + off];
      red[1][tid] += red[1][tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    beta[n] = red[0][0];
    tau[n] = red[1][0];
  }
}

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace lakde

using namespace lakde;

extern "C" {

int lakde_pack_records(const float *Linv, const float *Xc, const float *nu, int64_t N, int64_t D,
                       float *rec, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(Linv && Xc && nu && rec && N >= 0, "bad args");
  LAKDE_CHECK_ARG(D >= 1 && D <= 63, "D must be in [1, 63]");
  if (N == 0) return LAKDE_OK;
  hipLaunchKernelGGL(pack_records_kernel, dim3((unsigned)N), dim3(64), 0, as_stream(stream), Linv,
                     Xc, nu, (int)D, rec);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

size_t lakde_pred_backward_workspace(int64_t N, int64_t M, int64_t D) {
  int64_t per, ns;
  pred_geometry(N, M, &per, &ns);
  return lakde_pred_workspace(N, M) + al256((size_t)ns * M * D * 4) + al256((size_t)M * 4);
}

int lakde_pred_backward(const float *rec, int64_t N, const float *Y, int64_t M, int64_t D,
                        const float *out, const float *g, float *dY, float *S, float *s1,
                        float *beta, float *tau, void *workspace, size_t workspace_bytes,
                        lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 1 && D <= 63, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(rec && Y && out && g && dY && S && s1 && beta && tau && workspace, "null pointer");
  LAKDE_CHECK_ARG(workspace_bytes >= lakde_pred_backward_workspace(N, M, D), "workspace too small");
  if (M == 0 || N == 0) return LAKDE_OK;
  PParams p;
  const size_t base = lakde_pred_workspace(N, M);
  int rc = pred_fill_params(rec, N, M, (float *)out, workspace, base, &p);
  if (rc) return rc;
  p.Y = Y;
  float *part = (float *)((char *)workspace + base);
  float *lse = (float *)((char *)workspace + base + al256((size_t)p.nsplit * M * D * 4));
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(lse_from_out_kernel, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, out,
                     M, (float)std::log((double)N), lse);
  LAKDE_LAUNCHED();
  rc = dispatch_bwd_y(D, p, lse, g, part, s);
  if (rc) { set_error("lakde_pred_backward: dY launch failed"); return rc; }
  const int64_t MD = M * D;
  hipLaunchKernelGGL(sum_splits_kernel, dim3((unsigned)ceil_div(MD, 256)), dim3(256), 0, s, part,
                     p.nsplit, MD, dY);
  LAKDE_LAUNCHED();
  hipLaunchKernelGGL(bwd_kernel_stats, dim3((unsigned)N), dim3(256), 0, s, rec, (int)D, Y, M, lse,
                     g, S, s1, beta, tau);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

}  // extern "C"
