// k-nearest-neighbour graph (lakde/utils.py:77-116, _compute_knn_graph_generic
// with metric "l2"): for every point m the k points n with the smallest
// squared Euclidean distance, sorted ascending; equal distances keep the
// smaller index first, and n == m is left out unless include_self.  The
// reference ranks torch.cdist distances with torch.topk, whose order among
// ties is implementation-defined; here the order is (distance, index).
//
// Scan kernel: one lane per query point m (its coordinates in VGPRs), the
// candidates n walked in index order with wave-uniform (scalar) loads, the
// distance sum_j (x_m[j] - x_n[j])^2 accumulated in FP32 in j order.  A
// candidate closer than the lane's current k-th distance tau is queued in
// LDS; when some lane's queue is nearly full the wave merges every lane's
// sorted queue into that lane's sorted list (row-major, k entries, in the
// workspace) from the back, which stops as soon as the queue is used up.
// Candidates arrive in increasing index order, so a candidate whose
// distance equals tau would rank after the list's last entry: the strict
// test d < tau is the exact admission rule.
//
// Small grids split the candidate range over blockIdx.y; each split keeps
// its own top-k list and the merge kernel combines the splits of a query
// (their index ranges are disjoint) and writes the graph.
#include "common.hpp"
#include "qf.hpp"

namespace lakde {

struct KnnEntry {
  float d;
  int32_t i;
};

constexpr int kKnnQueue = 16;
constexpr int kKnnMaxSplit = 16;

__device__ __forceinline__ bool knn_after(float da, int32_t ia, float db, int32_t ib) {
  return da > db || (da == db && ia > ib);
}

struct KnnState {
  int lsize;  // entries in the lane's list
  float tau;  // admission bound: the k-th distance, +inf while the list is short
};

// Merge this lane's queue (cnt entries in q_d/q_i[.][tid], in arrival
// order) into its sorted list L of st.lsize entries, keeping the k smallest.
__device__ __noinline__ KnnState knn_flush(float (*q_d)[256], int32_t (*q_i)[256], int tid,
                                           int cnt, KnnEntry *__restrict__ L, int k, KnnState st) {
  if (cnt == 0) return st;
  const int lsize = st.lsize;
  // insertion sort of the queue by (distance, index)
  for (int a = 1; a < cnt; ++a) {
    const float kd = q_d[a][tid];
    const int32_t ki = q_i[a][tid];
    int b = a - 1;
    while (b >= 0 && knn_after(q_d[b][tid], q_i[b][tid], kd, ki)) {
      q_d[b + 1][tid] = q_d[b][tid];
      q_i[b + 1][tid] = q_i[b][tid];
      --b;
    }
    q_d[b + 1][tid] = kd;
    q_i[b + 1][tid] = ki;
  }
  const int total = lsize + cnt;
  const int keep = total < k ? total : k;
  int i = lsize - 1, j = cnt - 1;
  // list entries L[i], L[i-1], L[i-2], L[i-3] are held in registers and the
  // next one is loaded four steps before it is compared, so the merge does
  // not wait for a load per step.  Every prefetched index is below every
  // index written (out > i), so the prefetches read the list as it was.
  const KnnEntry pad{0.f, 0};
  KnnEntry w0 = i >= 0 ? L[i] : pad, w1 = i >= 1 ? L[i - 1] : pad;
  KnnEntry w2 = i >= 2 ? L[i - 2] : pad, w3 = i >= 3 ? L[i - 3] : pad;
  auto pop = [&]() {
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = i >= 4 ? L[i - 4] : pad;
    --i;
  };
  // drop the total - keep greatest entries of the union
  for (int t = total - keep; t > 0; --t) {
    if (j >= 0 && (i < 0 || knn_after(q_d[j][tid], q_i[j][tid], w0.d, w0.i)))
      --j;
    else
      pop();
  }
  // fill from the back; out == i + j + 1 throughout, so no unread list
  // entry is overwritten and the entries below out are in place once the
  // queue is exhausted.  The entry that ends at keep - 1 is the new bound.
  float last = (j < 0 && i >= 0) ? w0.d : 0.f;
  for (int out = keep - 1; j >= 0; --out) {
    KnnEntry e;
    if (i >= 0 && knn_after(w0.d, w0.i, q_d[j][tid], q_i[j][tid])) {
      e = w0;
      pop();
    } else {
      e = KnnEntry{q_d[j][tid], q_i[j][tid]};
      --j;
    }
    L[out] = e;
    if (out == keep - 1) last = e.d;
  }
  return KnnState{keep, keep == k ? last : __builtin_huge_valf()};
}

template <int D>
__device__ __forceinline__ float knn_dist(const float (&xm)[D], const float *__restrict__ xn) {
  float d = 0.f;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const float t = xm[j] - xn[j];
    d = fmaf(t, t, d);
  }
  return d;
}

template <int D>
__global__ void __launch_bounds__(256)
knn_scan_kernel(const float *__restrict__ X, int64_t N, int k, int include_self, int64_t per_split,
                KnnEntry *__restrict__ lists, int32_t *__restrict__ sizes) {
  constexpr int U = D <= 8 ? 4 : (D <= 24 ? 2 : 1);  // candidates per step
  __shared__ float q_d[kKnnQueue][256];
  __shared__ int32_t q_i[kKnnQueue][256];
  const int tid = threadIdx.x;
  const int64_t m = (int64_t)blockIdx.x * 256 + tid;
  const bool valid = m < N;
  const int64_t split = blockIdx.y;
  const int64_t nb = split * per_split;
  const int64_t ne = nb + per_split < N ? nb + per_split : N;
  float xm[D];
#pragma unroll
  for (int j = 0; j < D; ++j) xm[j] = valid ? X[m * D + j] : 0.f;
  KnnEntry *L = lists + (split * N + (valid ? m : 0)) * k;
  const int64_t self = (valid && !include_self) ? m : -1;
  int cnt = 0;
  KnnState st{0, valid ? __builtin_huge_valf() : -__builtin_huge_valf()};

  int64_t n = nb;
  for (; n + U <= ne; n += U) {
    float d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = knn_dist<D>(xm, X + (n + u) * D);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (d[u] < st.tau && n + u != self) {
        q_d[cnt][tid] = d[u];
        q_i[cnt][tid] = (int32_t)(n + u);
        ++cnt;
      }
    }
    if (__any(cnt > kKnnQueue - U)) {
      st = knn_flush(q_d, q_i, tid, cnt, L, k, st);
      cnt = 0;
    }
  }
  for (; n < ne; ++n) {
    const float d = knn_dist<D>(xm, X + n * D);
    if (d < st.tau && n != self) {
      q_d[cnt][tid] = d;
      q_i[cnt][tid] = (int32_t)n;
      ++cnt;
    }
    if (__any(cnt == kKnnQueue)) {
      st = knn_flush(q_d, q_i, tid, cnt, L, k, st);
      cnt = 0;
    }
  }
  st = knn_flush(q_d, q_i, tid, cnt, L, k, st);
  if (valid) sizes[split * N + m] = st.lsize;
}

// Combine the S split lists of each query (disjoint index ranges, each
// sorted by (distance, index)) and write the first k entries.
__global__ void __launch_bounds__(256)
knn_merge_kernel(const KnnEntry *__restrict__ lists, const int32_t *__restrict__ sizes, int64_t N,
                 int k, int S, void *__restrict__ out, int out_dtype, float *__restrict__ dist) {
  __shared__ int32_t pos[kKnnMaxSplit][256];
  const int tid = threadIdx.x;
  const int64_t m = (int64_t)blockIdx.x * 256 + tid;
  if (m >= N) return;
  for (int s = 0; s < S; ++s) pos[s][tid] = 0;
  for (int j = 0; j < k; ++j) {
    int best = -1;
    KnnEntry be{0.f, 0};
    for (int s = 0; s < S; ++s) {
      const int p = pos[s][tid];
      if (p >= sizes[s * N + m]) continue;
      const KnnEntry e = lists[(s * N + m) * k + p];
      if (best < 0 || knn_after(be.d, be.i, e.d, e.i)) {
        best = s;
        be = e;
      }
    }
    // the host guarantees at least k candidates per query
    if (best >= 0) pos[best][tid] += 1;
    if (out_dtype == LAKDE_I64)
      reinterpret_cast<int64_t *>(out)[m * k + j] = be.i;
    else
      reinterpret_cast<int32_t *>(out)[m * k + j] = be.i;
    if (dist) dist[m * k + j] = sqrtf(be.d);
  }
}

static size_t knn_align256(size_t x) { return (x + 255) / 256 * 256; }

static void knn_geometry(int64_t N, int64_t k, int64_t *per_split, int64_t *nsplit) {
  // Splitting the candidates fills more of the chip when there are few
  // queries, but every split keeps its own top-k list, and the list work
  // grows with the number of lists: a split holds at least 32 k candidates.
  const int64_t nblk = ceil_div(N, 256);
  int64_t s = ceil_div(2048, nblk > 0 ? nblk : 1);
  if (s > kKnnMaxSplit) s = kKnnMaxSplit;
  const int64_t by_k = N / (32 * (k > 8 ? k : 8));
  if (s > by_k) s = by_k;
  if (s < 1) s = 1;
  *per_split = ceil_div(N, s);
  *nsplit = ceil_div(N, *per_split);
}

template <int D>
static int launch_knn_scan(const float *X, int64_t N, int k, int include_self, int64_t per_split,
                           int64_t nsplit, KnnEntry *lists, int32_t *sizes, hipStream_t s) {
  hipLaunchKernelGGL((knn_scan_kernel<D>), dim3((unsigned)ceil_div(N, 256), (unsigned)nsplit),
                     dim3(256), 0, s, X, N, k, include_self, per_split, lists, sizes);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

}  // namespace lakde

using namespace lakde;

extern "C" {

size_t lakde_knn_graph_workspace(int64_t N, int64_t k) {
  int64_t per, ns;
  knn_geometry(N, k, &per, &ns);
  return knn_align256((size_t)ns * N * k * sizeof(KnnEntry)) + knn_align256((size_t)ns * N * 4);
}

int lakde_knn_graph(const float *X, int64_t N, int64_t D, int64_t k, int include_self, void *out,
                    int out_dtype, float *dist, void *workspace, size_t workspace_bytes,
                    lakde_stream_t stream) {
  LAKDE_CHECK_ARG(D >= 1 && D <= 63, "D must be in [1, 63]");
  LAKDE_CHECK_ARG(N >= 0 && N <= INT32_MAX, "N must be in [0, 2^31 - 1]");
  LAKDE_CHECK_ARG(out_dtype == LAKDE_I32 || out_dtype == LAKDE_I64, "out dtype must be int32 or int64");
  if (N == 0 || k == 0) return LAKDE_OK;
  LAKDE_CHECK_ARG(k >= 1 && k <= N - (include_self ? 0 : 1),
                  "k must be in [1, N] (include_self) or [1, N - 1]");
  LAKDE_CHECK_ARG(X && out && workspace, "null pointer");
  LAKDE_CHECK_ARG(workspace_bytes >= lakde_knn_graph_workspace(N, k), "workspace too small");
  int64_t per, ns;
  knn_geometry(N, k, &per, &ns);
  KnnEntry *lists = (KnnEntry *)workspace;
  int32_t *sizes = (int32_t *)((char *)workspace + knn_align256((size_t)ns * N * k * sizeof(KnnEntry)));
  hipStream_t s = as_stream(stream);
  int rc = LAKDE_EINVAL;
  switch (D) {
#define CASE(d) \
  case d: rc = launch_knn_scan<d>(X, N, (int)k, include_self, per, ns, lists, sizes, s); break;
    LAKDE_FOR_EACH_D(CASE)
#undef CASE
  }
  if (rc) return rc;
  hipLaunchKernelGGL(knn_merge_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, lists,
                     sizes, N, (int)k, (int)ns, out, out_dtype, dist);
  LAKDE_LAUNCHED();
  return LAKDE_OK;
}

}  // extern "C"
