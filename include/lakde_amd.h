/*
 * lakde_amd -- C ABI of the MI355X (gfx950) kernels behind lakde's hot path.
 *
 * Plain pointers, sizes and strides; no framework types.  Every entry point
 * enqueues work on `stream` (a hipStream_t; NULL = the null stream), never
 * allocates device memory, and returns a status code (LAKDE_OK on success).
 * Only lakde_partial_step synchronises the host (once, on `stream`).  lakde_last_error() gives a message for the
 * most recent failure on the calling thread.
 *
 * Device pointers are expected for every array argument.  Arguments are
 * validated on the host (sizes, dtype codes, null pointers); values inside
 * device arrays (indices in range, sorted rows) are the caller's contract,
 * exactly as for the reference CUDA ops.
 *
 * Reference interfaces replaced (reference repository paths):
 *   lakde_ind2ptr        <- torch.ops.lakde.ind2ptr_cuda
 *                           csrc/cuda/convert_cuda.h:27, lakde/kernels.py:93-100
 *   lakde_w_inv          <- torch.ops.lakde.W_inv_cuda
 *                           csrc/cuda/W_inv_cuda.h:5-7, lakde/kernels.py:102-113
 *   lakde_chol_quad_form <- torch.ops.lakde.chol_quad_form_cuda
 *                           csrc/cuda/chol_quad_form.h:5-6, lakde/kernels.py:115-120
 * The fused entry points below have no single reference op; each names the
 * reference Python code whose device work it performs.
 */
#ifndef LAKDE_AMD_H
#define LAKDE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *lakde_stream_t; /* hipStream_t */

/* An order-independent column sum: a 128-bit fixed-point integer (value *
 * 2^80 in hi:lo; the top bit of hi flags NaN).  The kernels add values to it
 * with integer atomics, so the sum does not depend on the order the adds
 * arrive in and a fit is reproducible run to run; a value added and later
 * removed cancels exactly.  Arrays of them must read zero before first use
 * (device memory cleared to zero). */
typedef struct {
  uint64_t lo, hi;
} lakde_fx128;

/* status codes */
#define LAKDE_OK 0
#define LAKDE_EINVAL 1      /* bad argument (size, stride, null pointer) */
#define LAKDE_EDTYPE 2      /* unsupported dtype combination */
#define LAKDE_EHIP 3        /* a HIP runtime call failed */
#define LAKDE_ECAPACITY 4   /* lakde_partial_step: the block's entries exceed cap */
#define LAKDE_ESTABILIZE 5  /* lakde_partial_step: a stabilisation did not converge */

/* dtype codes */
#define LAKDE_I32 0
#define LAKDE_I64 1
#define LAKDE_F32 2
#define LAKDE_F64 3
#define LAKDE_F16 4  /* storage only: the reference ops' float16 form computes in float32 */

/* E-step modes (lakde_estep) */
#define LAKDE_ESTEP_ACCUM 0  /* c[m] = logaddexp(c[m], LSE_n log rho_nm)         */
#define LAKDE_ESTEP_FIXED 1  /* responsibilities with c fixed (no renormalise)   */
#define LAKDE_ESTEP_UPDATE 2 /* incremental renormalise, then responsibilities   */

/* prior modes (lakde_posterior) */
#define LAKDE_PRIOR_KNN 0    /* per-point kNN covariance, factored with jitter   */
#define LAKDE_PRIOR_SHARED 1 /* one shared D x D prior covariance                */
#define LAKDE_PRIOR_KNN_CACHED 2 /* kNN prior read from lakde_knn_prior_cache's cache */

int lakde_version(void);
const char *lakde_last_error(void);

/* ------------------------------------------------------------------ *
 * Reference ops                                                       *
 * ------------------------------------------------------------------ */

/* rowptr[i] = #{e < nnz : ind[e] < i} for i = 0..M, ind sorted ascending.
 * ind: int32/int64 (ind_dtype); out: M+1 entries of out_dtype (int32/int64).
 * Replaces ind2ptr_cuda (csrc/cuda/convert_cuda.h:27; lakde/kernels.py:93-100). */
int lakde_ind2ptr(const void *ind, int ind_dtype, int64_t nnz, int64_t M,
                  void *out, int out_dtype, lakde_stream_t stream);

/* out[n] = sum_{e in [rowptr[n], rowptr[n+1])} value[e] d_e d_e^T,
 * d_e = Xm[col[e]] - Xn[n]; out is N x D x D contiguous, both triangles.
 * rowptr (N+1) and col are int32/int64; value, Xn (N x D), Xm (M x D) and out
 * are fp_dtype (F16/F32/F64; F16 values are read into and summed in float32,
 * the result rounded once to float16), all contiguous.
 * Replaces W_inv_cuda (csrc/cuda/W_inv_cuda.h:5-7, its dispatch over float16,
 * float32 and float64 at csrc/cuda/W_inv_cuda.cu:102-103; lakde/kernels.py:102-113). */
int lakde_w_inv(const void *rowptr, int rowptr_dtype, const void *col, int col_dtype,
                const void *value, const void *Xn, const void *Xm, int64_t N, int64_t M,
                int64_t D, int fp_dtype, void *out, lakde_stream_t stream);

/* out[n, m] = || W[n]^T (Xm[m] - Xn[n]) ||^2 with W[n] upper triangular
 * (entries below the diagonal are ignored).  Strides are in elements and may
 * be zero or negative-free arbitrary (stride-0 broadcasts are allowed).
 * fp_dtype F16/F32/F64 (F16: float32 arithmetic, the result rounded to
 * float16).  Replaces chol_quad_form_cuda (csrc/cuda/chol_quad_form.h:5-6,
 * dispatched over the same three types at csrc/cuda/chol_quad_form.cu:105-106;
 * lakde/kernels.py:115-120). */
int lakde_chol_quad_form(const void *W, int64_t sW0, int64_t sW1, int64_t sW2,
                         const void *Xn, int64_t sXn0, int64_t sXn1,
                         const void *Xm, int64_t sXm0, int64_t sXm1,
                         void *out, int64_t so0, int64_t so1,
                         int64_t N, int64_t M, int64_t D, int fp_dtype,
                         lakde_stream_t stream);

/* ------------------------------------------------------------------ *
 * Fused hot path (float32)                                            *
 * ------------------------------------------------------------------ */

/* Floats per packed kernel record: L^{-1} (lower, packed row-major),
 * the centre x_n, and four per-kernel constants; padded to a multiple of 4. */
int64_t lakde_record_floats(int64_t D);

/* kNN prior covariance of rows [row0, row0+nb):
 * out[n] = (sum_k (X[g[n,k]] - X[row0+n]) (...)^T) / K.
 * graph: nb x K (int32/int64), rows already offset to the block.
 * Performs batched_knn_cov (lakde/utils.py:40-65, self_center=True). */
int lakde_knn_prior_cov(const float *X, int64_t N, int64_t D, int64_t row0, int64_t nb,
                        const void *graph, int graph_dtype, int64_t K, float *out,
                        lakde_stream_t stream);

/* S[n] as lakde_w_inv on one block's CSR plus nu[n] = nu0 + sum_e value[e]
 * (lakde/local_adaptive_kde.py:66-81). rowptr/col int32/int64.  With
 * colscale (N floats, or NULL) the responsibility of entry e is
 * value[e] * colscale[col[e]] (stored values plus per-column factors, see
 * lakde_rebalance_lazy). */
int lakde_posterior_contribs(const void *rowptr, int rowptr_dtype, const void *col,
                             int col_dtype, const float *value, const float *colscale,
                             const float *Xn, const float *Xm, int64_t nb, int64_t M, int64_t D,
                             float nu0, float *S, float *nu, lakde_stream_t stream);

/* Batched posterior factorisation (lakde/local_adaptive_kde.py:83-96,
 * lakde/linalg_util.py:30-37,49-52,109-170, lakde/local_knn_kde.py:195-207,220-222,
 * lakde/local_hierarchical_kde.py:269-271,289-291): one wavefront per kernel.
 *   kNN prior:   P_n = chol(prior[n] + j0_n I) with persistent jitter j0;
 *                prior_j = prior[n] + j0_n I; T = P_n.
 *   shared:      prior_j = prior (one D x D matrix); T = prior.
 *   W_inv = prior_j * (nu0 * a_n / b_n) + S[n];  L = chol(W_inv + j_n I), with
 *   the reference's jitter schedule decided in FP32; jittered factors are
 *   refactored and inverted in FP64.  Optionally the tau update
 *   b_out = b0 + nu0 * trace / 2 (lakde/local_knn_kde.py:236,
 *   lakde/local_hierarchical_kde.py:290) after E[tau] has been read.  Outputs: Linv = L^{-1} (lower; the
 *   reference's W_triu is Linv^T), logdet_W (FP64-accumulated), trace =
 *   nu_n * ||T^T Linv^T||_F^2, records for the E-step / predictive kernels.
 * *status (device int) is set to 1 when a stabilisation does not converge
 * within 32 doublings (the reference raises RuntimeError). */
typedef struct {
  int64_t nb, D;
  int prior_mode;
  const float *prior;     /* kNN: nb x D x D; shared: D x D; cached: nb cache entries */
  float *prior_jitter;    /* kNN: nb, in/out; shared: NULL */
  const float *S;         /* nb x D x D */
  const float *nu;        /* nb */
  const float *a, *b;     /* nb (E[tau] = a / b) */
  float nu0;
  float *W_jitter;        /* nb, in/out */
  const float *Xn;        /* nb x D */
  float *Linv;            /* nb x D x D out */
  float *prior_tril;      /* kNN: nb x D x D out; shared: NULL */
  float *logdet_W;        /* nb out */
  float *logdet_prior;    /* kNN: nb out; shared: NULL */
  float *trace;           /* nb out */
  float *rec;             /* nb x lakde_record_floats(D) out */
  int *status;            /* device int */
  float b0;               /* with b_out: b_out[n] = b0 + nu0 * trace[n] / 2 */
  float *b_out;           /* nb out (may alias b), or NULL */
} lakde_posterior_args;

int lakde_posterior(const lakde_posterior_args *args, lakde_stream_t stream);

/* The kNN prior's jittered factor, cached for a fit: X and the graph do not
 * change during a fit, so the prior covariance of every point and, with its
 * persistent jitter, its factor are those of the first computation
 * (lakde/local_knn_kde.py:195-207 recomputes both every partial step and
 * every prediction).  For rows [row0, row0+nb): lakde_knn_prior_cov into
 * prior_scratch (nb x D x D), then each factored with the stabilised jitter
 * schedule (prior_jitter in/out, logdet_prior out, *status = 1 when it does
 * not converge), and cache entry n (lakde_prior_cache_floats(D) floats) =
 * prior + j0 I then its Cholesky factor, both packed lower row-major.  A
 * posterior with prior_mode LAKDE_PRIOR_KNN_CACHED reads prior = the block's
 * entries and leaves prior_jitter, prior_tril and logdet_prior alone (D >= 7;
 * the fused small-D posterior computes its prior in place). */
int64_t lakde_prior_cache_floats(int64_t D);
int lakde_knn_prior_cache(const float *X, int64_t N, int64_t D, int64_t row0, int64_t nb,
                          const void *graph, int graph_dtype, int64_t K, float *prior_scratch,
                          float *prior_jitter, float *logdet_prior, float *cache, int32_t *status,
                          lakde_stream_t stream);

/* One block's posterior from one call: for a kNN prior (knn_graph = the
 * block's nb x knn_k graph rows, int32/int64) lakde_knn_prior_cov into
 * post->prior, then lakde_posterior_contribs of the block's CSR (rowptr nb+1,
 * col, val; int32 indices; optional colscale) into post->S, post->nu with
 * nu0 = post->nu0, then lakde_posterior(post).  A shared prior passes
 * knn_graph = NULL.  For D <= 6 one fused kernel does all three (a wavefront
 * per kernel point; post->prior and post->S are then not written).
 * (lakde/local_knn_kde.py:195-207 calc_prior_cov,
 * lakde/local_adaptive_kde.py:66-96 calc_posterior_contribs + calc_posterior) */
int lakde_block_posterior(const lakde_posterior_args *post, const float *X, int64_t N,
                          int64_t row0, const int32_t *rowptr, const int32_t *col,
                          const float *val, const float *colscale, const void *knn_graph,
                          int knn_graph_dtype, int64_t knn_k, lakde_stream_t stream);

/* Column sums of one block's responsibilities: sums[m] = sum_{e: col[e]=m} val[e]
 * (written, not accumulated), summed order-independently in ws (N
 * lakde_fx128, zero on entry and left zero).  (lakde/local_adaptive_kde.py:120) */
int lakde_colsum(const int32_t *col, const float *val, const int32_t *nnz_dev, int64_t N,
                 float *sums, lakde_fx128 *ws, lakde_stream_t stream);

/* Fused E-step of one kernel block over all N points
 * (lakde/local_adaptive_kde.py:36-44,98-153, lakde/local_knn_kde.py:157-166):
 * quadratic forms from the records, log rho epilogue, diagonal exclusion,
 * streaming log-sum-exp over the block's kernels, incremental renormalisation
 * (UPDATE), exp / threshold and compaction to canonical CSR (row ascending,
 * column ascending).  Responsibilities r > threshold are kept (threshold 0:
 * r != 0).  Output arrays hold `cap` entries; *nnz_dev receives the true
 * count and nothing is written past cap (the caller grows and re-runs
 * lakde_estep_compact).  Workspace: lakde_estep_workspace() bytes, zero-filled
 * before its first use; it holds counters that every ESTEP_FIXED/UPDATE
 * launch of lakde_estep_main accumulates and lakde_estep_finish clears, so a
 * main launch in those modes must be followed by lakde_estep_finish on the
 * same workspace. */
typedef struct {
  int64_t N, D, nb, row0;
  const float *X;          /* N x D */
  const float *rec;        /* nb records */
  int mode;
  float threshold;
  const float *c_old;      /* N (UPDATE: previous log normalisers) */
  const float *old_colsum; /* N (UPDATE: column sums of the block's old values) */
  float *c;                /* N: ACCUM in/out, FIXED in, UPDATE out */
  int32_t *rowptr;         /* nb + 1 out */
  int32_t *out_row, *out_col;
  float *out_val;
  int64_t cap;
  int32_t *nnz_dev;        /* device int out */
  void *workspace;
  size_t workspace_bytes;
  /* optional (UPDATE): the unsafe-step test of lakde_rebalance_lazy_check
   * fused into the main kernel: *lazy_unsafe |= 1 for a column whose
   * lazy_G[m] clamp(exp(c_old - c)) leaves [2^-40, 2^40] (the caller clears
   * *lazy_unsafe beforehand); NULL: no test */
  const float *lazy_G;
  int32_t *lazy_unsafe;
  /* optional (1 <= D <= 16): column order for pair skipping.  col_perm (N
   * int32) lists the columns so that every 64 consecutive entries form a
   * spatially compact chunk (e.g. a k-d ordering with 64-point leaves);
   * chunk_box holds each chunk's bounding box (ceil(N/64) x 2 x D floats:
   * lo then hi).  An UPDATE launch then takes its columns chunk by chunk and
   * skips every kernel whose log rho is provably more than
   * max(30 + log(nb/512), 8 - log threshold) nats (110 for threshold 0)
   * below the chunk's smallest lower bound of the new normaliser on the
   * whole box.  Such pairs change no normaliser at FP32 resolution and are
   * never kept, so the outputs are those of the full evaluation.  With
   * col_perm set, the workspace's masks must be zero on entry to a launch in
   * any mode: a zero-filled workspace whose every compaction ran with
   * col_perm set satisfies this (the compaction clears the masks it reads).
   * NULL: columns in index order, no skipping. */
  const int32_t *col_perm;
  const float *chunk_box;
  /* optional (with lazy_unsafe, UPDATE): N floats, zero on entry.  A column
   * whose lazy_G[m] clamp(exp(c_old - c)) falls below 2^-40 but stays above
   * 2^-100, with clamp(exp(c_old - c)) >= 2^-100, gets that factor here and
   * *lazy_unsafe |= 2 (re-basing the column, see lakde_step_args) instead of
   * *lazy_unsafe |= 1.  NULL: every column outside the window sets 1. */
  float *lazy_factor;
} lakde_estep_args;

size_t lakde_estep_workspace(int64_t N, int64_t nb);
int lakde_estep(const lakde_estep_args *args, lakde_stream_t stream);
/* lakde_estep in two halves: the quad-form / LSE / ballot kernel, then the
 * row scans and the compaction (lets a caller time the main kernel alone). */
int lakde_estep_main(const lakde_estep_args *args, lakde_stream_t stream);
int lakde_estep_finish(const lakde_estep_args *args, lakde_stream_t stream);
/* Re-run only the compaction of the last lakde_estep (same args, larger cap). */
int lakde_estep_compact(const lakde_estep_args *args, lakde_stream_t stream);

/* rebalance_responsibilities_ (lakde/local_adaptive_kde.py:156-181) over a
 * block arena: block b's entries live at [b*cap, b*cap + nnz[b]).  Blocks other
 * than `exclude_block` are scaled by clamp(exp(c_prev - c), tiny, max)[col];
 * column sums s are clamped to eps; c += log s; every value is scaled by
 * clamp(1/s, tiny, max)[col].  c_prev may alias c.  The column sums are
 * order-independent: s = coeff (sum of the other blocks' old values) + (sum
 * of the excluded block's values), each a lakde_fx128 sum.  ws:
 * lakde_rebalance_workspace(N) bytes, zero on first use and left so. */
size_t lakde_rebalance_workspace(int64_t N);
int lakde_rebalance(int64_t N, int64_t num_blocks, int64_t cap, int64_t exclude_block,
                    const int32_t *nnz, const int32_t *cols, float *vals,
                    const float *c_prev, float *c, void *ws, lakde_stream_t stream);

/* The same rebalance with the responsibilities stored unscaled (lazy
 * rebalance): the value of entry e is val[e] * G[col[e]] with per-column
 * factors G (float) and stored column totals U.  A partial step on block i
 * then touches only block i's entries and O(N) column state instead of
 * every block:
 *   lakde_rebalance_lazy_begin (before lakde_estep, ESTEP_UPDATE):
 *     old_colsum[m] = sum over block i of val*G (the true old column sums
 *     the E-step reads; written), U_old[m] += sum of block i's stored
 *     values, *unsafe = 0;
 *   lakde_rebalance_lazy_check (after the E-step wrote c):
 *     *unsafe = 1 if any G[m] clamp(exp(c_prev - c)) lies outside
 *     [2^-40, 2^40] or is NaN; the stored form cannot represent that step
 *     and the caller must instead fold (lakde_rebalance_fold), run
 *     lakde_rebalance, fold again and clear U_old;
 *   lakde_rebalance_lazy (otherwise; block i now holds true values r):
 *     coeff = clamp(exp(c_prev - c), tiny, max), val = r / (G coeff),
 *     s = max(coeff G (U - U_old) + sum r, eps), c += log s,
 *     G *= coeff clamp(1/s, tiny, max), U += sum val - U_old,
 *     and U_old is cleared again.
 * U, U_old: N lakde_fx128 each (order-independent sums: U - U_old is exact,
 * so a column whose entries all lie in block i reads exactly zero); U_old is
 * zero between steps; ws is 3*N lakde_fx128, zero on entry and left zero.
 * With col, val, nnz_dev NULL no block was updated (the initial
 * normalisation, c_prev == c).  c_prev may alias c. */
int lakde_rebalance_lazy_begin(int64_t N, const int32_t *col, const float *val,
                               const int32_t *nnz_dev, const float *G, float *old_colsum,
                               lakde_fx128 *U_old, lakde_fx128 *ws, int32_t *unsafe,
                               lakde_stream_t stream);
int lakde_rebalance_lazy_check(int64_t N, const float *c_prev, const float *c, const float *G,
                               int32_t *unsafe, lakde_stream_t stream);
int lakde_rebalance_lazy(int64_t N, const int32_t *col, float *val, const int32_t *nnz_dev,
                         const float *c_prev, float *c, float *G, lakde_fx128 *U,
                         lakde_fx128 *U_old, lakde_fx128 *ws, lakde_stream_t stream);

/* Fold the column factors into the stored values of every block of the
 * arena (val *= G[col]), then G = 1 and U = the exact stored column sums.
 * Bounds the drift of U; afterwards the arena holds true values. */
int lakde_rebalance_fold(int64_t N, int64_t num_blocks, int64_t cap, const int32_t *nnz,
                         const int32_t *cols, float *vals, float *G, lakde_fx128 *U,
                         lakde_stream_t stream);

/* ------------------------------------------------------------------ *
 * Partial-step driver                                                 *
 * ------------------------------------------------------------------ */

/* One partial step of kernel block `block` (lakde/local_knn_kde.py:209-239 and
 * lakde/local_hierarchical_kde.py:273-311 up to the shared-prior update, i.e.
 * lakde/local_adaptive_kde.py:66-181): every launch of the step from one
 * call, then one copy of the arena counts and flags into host_flags and one
 * stream synchronisation, then the rebalance.
 *
 *   knn_graph != NULL: lakde_knn_prior_cov into post.prior first;
 *   lakde_posterior_contribs of the block's current entries (estep.rowptr,
 *     estep.out_col, estep.out_val; colscale G) into post.S, post.nu;
 *   lakde_posterior(post);
 *   G != NULL (lazy): [fold_first: lakde_rebalance_fold] and
 *     lakde_rebalance_lazy_begin; else lakde_colsum into estep.old_colsum;
 *   bf16_frag != NULL: lakde_pack_frag_bf16 of estep.rec and
 *     lakde_estep_main_bf16, else lakde_estep_main (ev_main[0], ev_main[1]:
 *     optional hipEvent_t recorded around it); lakde_estep_finish;
 *   the row counts of the E-step; lazy: lakde_rebalance_lazy_check, with
 *     the re-basing test (estep.lazy_factor in rebal_ws): a step whose only
 *     out-of-window columns are re-basable (flag bit 1, bit 0 clear) has
 *     those columns' stored values in the other blocks multiplied by their
 *     factor f = G coeff, their totals recomputed and G divided by f on the
 *     device before the compaction, and continues as a lazy step;
 *   host_flags[0 .. num_blocks+1] <- nnz_all (per-block counts, status,
 *     unsafe flag) [record ev_flags]; the compaction and the guarded lazy
 *     rebalance; wait for ev_flags (or the whole stream without it), so the
 *     compaction and rebalance overlap the host's next launches;
 *   status set: LAKDE_ESTABILIZE; host_flags[block] > cap: LAKDE_ECAPACITY
 *     (the caller grows the arena, updates the pointers and calls
 *     lakde_partial_step_resume, which compacts again and rebalances);
 *   otherwise the rebalance: lazy (lakde_rebalance_lazy, enqueued before the
 *   copy with a device-side guard that skips it for an unsafe (flag bit 0)
 *   or overflowing step; for an unsafe step the exact rebalance of the true
 *   values, the other blocks folded on the way, after the synchronisation)
 *   or lakde_rebalance(exclude block).
 *
 * estep must be an ESTEP_UPDATE step whose count, status and entry arrays are
 * the block's slots of the arena: estep.nnz_dev == nnz_all + block,
 * post.status == nnz_all + num_blocks, the unsafe flag is
 * nnz_all[num_blocks + 1], estep.out_col/out_val == cols_all/vals_all +
 * block * cap.  host_flags: num_blocks + 2 ints of pinned host memory. */
typedef struct {
  lakde_posterior_args post;
  lakde_estep_args estep;
  const void *knn_graph;   /* nb x K rows of the block, or NULL */
  int knn_graph_dtype;
  int64_t knn_k;
  float nu0;
  int64_t block, num_blocks, cap;
  int32_t *nnz_all;        /* num_blocks + 2 */
  int32_t *cols_all;       /* num_blocks x cap */
  float *vals_all;         /* num_blocks x cap */
  float *G;                /* N, lazy rebalance factors, or NULL (exact) */
  lakde_fx128 *U, *U_old;  /* N each (lazy) */
  lakde_fx128 *lazy_ws;    /* 3 N (lazy) */
  void *rebal_ws;          /* lakde_rebalance_workspace(N) bytes, zero-filled once: exact
                              rebalance, the exact path's old column sums, the lazy
                              fallback and the lazy re-basing state */
  int fold_first;          /* lazy: fold G into the values before the step */
  void *bf16_frag;         /* nb * lakde_bf16_frag_bytes(D) bytes, or NULL */
  int32_t *host_flags;     /* pinned host, num_blocks + 2 */
  void *ev_main[2];
  void *ev_flags;          /* optional hipEvent_t: the wait is for the flags copy only */
  /* optional, from lakde_flag_words_alloc: the row-pointer kernel writes this
   * block's entry count and the status / unsafe flags there (host-coherent
   * memory, sequence-tagged words) and the driver waits for them instead of a
   * copy and an event; it stores them into host_flags[block],
   * host_flags[num_blocks], host_flags[num_blocks + 1] */
  void *flag_words;
  /* optional (with flag_words; FP32, lazy rebalance, D <= 6): a device int
   * that the row-pointer kernel of every step sets to 1 when the step needs
   * the host (entries beyond cap, a failed stabilisation, an unsafe lazy
   * step) and to 0 otherwise.  lakde_partial_sweep then enqueues the next
   * step's posterior before it waits for this step's flags; that posterior
   * does nothing when the word is 1, and is enqueued again after the host
   * has handled the step.  NULL: no posterior is enqueued ahead. */
  int32_t *abort_flag;
  /* optional (with bf16_frag of lakde_screen_frag_bytes(D) per kernel):
   * the screening E-step (lakde_pack_screen + lakde_estep_main_screen) with
   * the points of lakde_screen_points and its origin; NULL: bf16_frag holds
   * lakde_pack_frag_bf16 blocks for lakde_estep_main_bf16 */
  const void *screen_xb;
  const float *screen_xnorm;
  const float *screen_origin;
  /* 1: bf16_frag holds lakde_m32_frag_bytes(D) per kernel and the step runs
   * the FP32 matrix-core E-step (lakde_pack_m32 + lakde_estep_main_m32,
   * with screen_xnorm and screen_origin; screen_xb may be NULL) */
  int fp32_mfma;
  /* optional (with the screening E-step): lakde_screen_split_workspace(N, nb)
   * bytes; the step runs lakde_estep_main_screen_split.  NULL: unsplit */
  void *screen_split_ws;
  size_t screen_split_bytes;
} lakde_step_args;

/* A 256-byte host-coherent, device-mapped buffer (hipHostMalloc,
 * hipHostMallocCoherent | hipHostMallocMapped) for lakde_step_args.flag_words:
 * words 0-2 receive the flags, word 3 is the driver's step counter, word 4
 * holds the device-side address; one per concurrently driven model.  No reference counterpart: it replaces the per-step device-to-host
 * copy of the flags and the event the host waits on. */
int lakde_flag_words_alloc(void **out);
int lakde_flag_words_free(void *words);

int lakde_partial_step(const lakde_step_args *args, lakde_stream_t stream);
int lakde_partial_step_resume(const lakde_step_args *args, lakde_stream_t stream);

/* Partial steps of blocks first .. first + count - 1 in order (one VI
 * iteration when first = 0 and count = num_blocks; lakde/kdes.py:64-67):
 * steps[b] are the arguments of block b.  Step b reads the normalisers
 * c[*cur] and writes c[1 - *cur] (estep.c_old / estep.c are set here), after
 * which *cur flips.  Lazy rebalance: fold_first is set for block 0 when
 * *lazy_dirty, and *lazy_dirty is updated after every step.  Stops at the
 * first step that does not return LAKDE_OK, with *done = steps completed
 * (LAKDE_ECAPACITY: grow the arena, rebuild the arguments, finish that step
 * with lakde_partial_step_resume, flip *cur and continue after it). */
int lakde_partial_sweep(lakde_step_args *steps, int64_t first, int64_t count, float *const *c,
                        int *cur, int *lazy_dirty, int64_t *done, lakde_stream_t stream);

/* Streaming predictive density (lakde/local_adaptive_kde.py:200-223):
 * out[m] = -log N + LSE_n [coeff_n - (nu_n+1)/2 log1p(Q_nm)] over all N
 * kernel records.  Workspace: lakde_pred_workspace() bytes. */
size_t lakde_pred_workspace(int64_t N, int64_t M);
int lakde_pred_density(const float *rec, int64_t N, const float *Y, int64_t M, int64_t D,
                       float *out, void *workspace, size_t workspace_bytes,
                       lakde_stream_t stream);

/* lakde_pred_density without a workspace argument and without any device
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The per-split partial sums live in a library-owned
 * file-scope device array of LAKDE_PRED_NOSCRATCH_MAX_BYTES bytes, one per
 * GPU for the life of the process (its address is looked up once per device
 * on the first call there).  Calls whose partial sums need more than that
 * (lakde_pred_workspace(N, M) bytes: about 8 M bytes for M >= 4096, at most
 * 2 MiB below) fail with LAKDE_EINVAL.  Same kernels and results as
 * lakde_pred_density.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race. */
#define LAKDE_PRED_NOSCRATCH_MAX_BYTES ((size_t)16 << 20)
int lakde_pred_density_noscratch(const float *rec, int64_t N, const float *Y, int64_t M,
                                 int64_t D, float *out, lakde_stream_t stream);

/* k-nearest-neighbour graph for initialisation (lakde/utils.py:77-116,
 * _compute_knn_graph_generic with metric "l2"; called at init from
 * lakde/local_knn_kde.py:95 and lakde/local_hierarchical_kde.py:149).
 * X: N x D float32 row-major.  out: N x k (LAKDE_I32 or LAKDE_I64)
 * neighbour indices, each row sorted by squared distance (FP32, summed in
 * coordinate order), equal distances by index; n == m is excluded unless
 * include_self.  dist (optional, N x k): the Euclidean distances.
 * Workspace: lakde_knn_graph_workspace() bytes, no initial contents. */
size_t lakde_knn_graph_workspace(int64_t N, int64_t k);
int lakde_knn_graph(const float *X, int64_t N, int64_t D, int64_t k, int include_self, void *out,
                    int out_dtype, float *dist, void *workspace, size_t workspace_bytes,
                    lakde_stream_t stream);

/* ------------------------------------------------------------------ *
 * Precision form (8 <= D <= 63): BF16 matrix cores, split precision    *
 * ------------------------------------------------------------------ */

/* Bytes of one kernel's bf16 fragment block (Linv hi/lo, FP32 centre),
 * or -1 when D is outside [8, 63]. */
int64_t lakde_bf16_frag_bytes(int64_t D);

/* Fragment blocks of nb kernels from their FP32 records: Linv split into
 * bf16 hi + lo (zero padded) and the centre x_n in FP32. */
int lakde_pack_frag_bf16(const float *rec, int64_t nb, int64_t D, void *frag,
                         lakde_stream_t stream);

/* lakde_estep_main on the matrix cores: y = Linv (x_m - x_n) via bf16x3 MFMA
 * (hi*hi + hi*lo + lo*hi of the split factor and the split difference, FP32
 * accumulate).  Same outputs as lakde_estep_main; follow with
 * lakde_estep_finish, whose compaction recomputes every kept responsibility
 * exactly in FP32. */
int lakde_estep_main_bf16(const lakde_estep_args *args, const void *frag,
                          lakde_stream_t stream);

/* Screening E-step (the default precision form; 8 <= D <= 63).  Every pair
 * is screened on the BF16 matrix cores in the centre form
 * bf16(Linv_n) bf16(x_m - o) - Linv_n (x_n - o) (FP32 accumulate, the centre
 * term in FP32, o a fixed origin such as the data mean) with a rigorous
 * bound on its error; the pairs that can matter to the normaliser or the
 * keep test (about 0.5% at D = 63) are recomputed exactly in FP32, and the
 * log-sum-exp, normaliser update and keep test run on those exact values.
 * Outputs as lakde_estep_main (follow with lakde_estep_finish).
 *   lakde_screen_points: xb (lakde_screen_points_bytes(N, D) bytes) =
 *     bf16(X - o) zero padded per point, xnorm (2N floats): xnorm[m] =
 *     |x_m - o| and xnorm[N + m] = |fl(x_m - o) - bf16(x_m - o)| (the norm
 *     of the rounding residual that the error bound uses), both rounded
 *     up; once per fit (X and o fixed).
 *   lakde_pack_screen: per kernel of a block (lakde_screen_frag_bytes(D)
 *     bytes each) from its records and o.
 *   lakde_screen_knn_bound: out[m] = the largest exact log rho of point m
 *     under the kernels of its first J neighbours of graph (N x k, int32 or
 *     int64; J a power of two <= 64), a lower bound of the normaliser the
 *     first log-sum-exp passes of a fit accumulate.
 *   lbound (ACCUM only, may be NULL): such a bound; the passes drop the
 *     pairs far below it from the start instead of below the normaliser
 *     accumulated so far.
 * (lakde/local_adaptive_kde.py:36-44,98-153 as lakde_estep_main.) */
int64_t lakde_screen_frag_bytes(int64_t D);
size_t lakde_screen_points_bytes(int64_t N, int64_t D);
int lakde_screen_points(const float *X, int64_t N, int64_t D, const float *origin, void *xb,
                        float *xnorm, lakde_stream_t stream);
int lakde_pack_screen(const float *rec, int64_t nb, int64_t D, const float *origin, void *frag,
                      lakde_stream_t stream);
int lakde_screen_knn_bound(const float *rec, const float *X, int64_t N, int64_t D,
                           const void *graph, int graph_dtype, int64_t k, int64_t J, float *out,
                           lakde_stream_t stream);
int lakde_estep_main_screen(const lakde_estep_args *args, const void *frag, const void *xb,
                            const float *xnorm, const float *lbound, lakde_stream_t stream);
/* lakde_estep_main_screen in the split form (the same outputs up to float32
 * summation order): in UPDATE mode the block's kernels are divided into
 * ranges of about 128, screened by separate workgroups that queue their
 * surviving pairs, and each column's ranges are combined by a second kernel
 * (the exact FP32 values of the queued pairs of 64 columns in one list,
 * normaliser update, unsafe-step test, keep tests; a third kernel does the
 * keep pass for the columns whose keep candidates overflowed).  Other modes: as lakde_estep_main_screen
 * without a bound.  split_ws: lakde_screen_split_workspace(N, nb) bytes,
 * no contents needed on entry. */
size_t lakde_screen_split_workspace(int64_t N, int64_t nb);
int lakde_estep_main_screen_split(const lakde_estep_args *args, const void *frag, const void *xb,
                                  const float *xnorm, void *split_ws, size_t split_bytes,
                                  lakde_stream_t stream);

/* The plain FP32 E-step on the matrix cores (17 <= D <= 63): every pair's
 * quadratic form y = Linv_n (x_m - x_n) on v_mfma_f32_32x32x2_f32 (FP32
 * products and accumulation, the difference x_m - x_n formed in FP32), its
 * log rho in the log-sum-exp and the normaliser update; the keep test runs
 * on the exact values of the compaction for the pairs an error bound cannot
 * rule out.  Outputs as lakde_estep_main (follow with lakde_estep_finish).
 *   lakde_pack_m32: per kernel of a block (lakde_m32_frag_bytes(D) bytes
 *     each, 16-byte aligned) from its records and the origin o of
 *     lakde_screen_points (xnorm there: |x - o| per point).
 * (lakde/local_adaptive_kde.py:36-44,98-153 as lakde_estep_main.) */
int64_t lakde_m32_frag_bytes(int64_t D);
int lakde_pack_m32(const float *rec, int64_t nb, int64_t D, const float *origin, void *frag,
                   lakde_stream_t stream);
int lakde_estep_main_m32(const lakde_estep_args *args, const void *frag, const float *xnorm,
                         lakde_stream_t stream);

/* lakde_pred_density with the quadratic forms on the matrix cores
 * (frag: N fragment blocks; Y: M x D float32 test points). */
int lakde_pred_density_bf16(const float *rec, const void *frag, int64_t N, const float *Y,
                            int64_t M, int64_t D, float *out, void *workspace,
                            size_t workspace_bytes, lakde_stream_t stream);

/* The predictive density of the precision form with the screen of
 * lakde_estep_main_screen: every (kernel, test point) pair is bounded on the
 * BF16 matrix cores, pairs whose upper bound lies more than
 * log(kernels per split) + log(1e6) nats below a running lower bound of
 * their split's partial log-sum-exp are dropped (together below 1e-6 of it,
 * i.e. 1e-6 nats), the others are evaluated exactly in FP32 with the terms
 * of lakde_pred_density.  frag: lakde_pack_screen of all N kernels (16-byte
 * aligned); yb, ynorm: lakde_screen_points of Y with the origin frag was
 * packed with; workspace: lakde_pred_screen_workspace(N, M) bytes.
 * (lakde/local_adaptive_kde.py:200-223 as lakde_pred_density.) */
size_t lakde_pred_screen_workspace(int64_t N, int64_t M);
int lakde_pred_density_screen(const float *rec, const void *frag, int64_t N, const float *Y,
                              int64_t M, int64_t D, const void *yb, const float *ynorm,
                              float *out, void *workspace, size_t workspace_bytes,
                              lakde_stream_t stream);

/* ------------------------------------------------------------------ *
 * Differentiation form                                                *
 * ------------------------------------------------------------------ */

/* Records from explicit factors: Linv (N x D x D lower), centres Xc (N x D),
 * nu (N); the constants are derived as in lakde_posterior. */
int lakde_pack_records(const float *Linv, const float *Xc, const float *nu, int64_t N,
                       int64_t D, float *rec, lakde_stream_t stream);

/* Backward of lakde_pred_density (lakde/local_adaptive_kde.py:200-223 has
 * none; this is the gradient of its result).  Inputs: the records, Y, the
 * forward output `out` (M) and the upstream gradient g (M).  The mixture
 * weights w_nm are recomputed from out (nothing of size N x M is stored).
 * Outputs: dY (M x D) = d(sum_m g_m out_m)/dY, and per kernel
 *   S (N x D x D) = sum_m a_nm d d^T,  s1 (N x D) = sum_m a_nm d,
 *   beta (N) = sum_m g_m w_nm,  tau (N) = sum_m g_m w_nm log1p(Q_nm),
 * with d = y_m - x_n and a_nm = -g_m w_nm (nu_n + 1) / 2 / (1 + Q_nm), from
 * which d/dLinv = tril(2 Linv S) + diag(beta / diag(Linv)),
 * d/dXc = -2 Linv^T Linv s1 and d/dnu = (beta (psi(h) - psi(h - D/2)) - tau) / 2,
 * h = (nu + 1) / 2. */
size_t lakde_pred_backward_workspace(int64_t N, int64_t M, int64_t D);
int lakde_pred_backward(const float *rec, int64_t N, const float *Y, int64_t M, int64_t D,
                        const float *out, const float *g, float *dY, float *S, float *s1,
                        float *beta, float *tau, void *workspace, size_t workspace_bytes,
                        lakde_stream_t stream);

/* fill helper (device memset replacement that is safe under graph capture) */
int lakde_fill_f32(float *p, int64_t n, float v, lakde_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LAKDE_AMD_H */
