// Host-side driver of one partial step (block b): every launch of the step
// from one C call, with the single host synchronisation it needs (the new
// block's entry count, the stabilisation status and the lazy rebalance's
// unsafe-step flag, copied together into pinned host memory).
//
// Device work, in stream order (lakde/local_knn_kde.py:209-239,
// lakde/local_adaptive_kde.py:66-181):
//   [lazy, first block of a sweep: fold]        lakde_rebalance_fold
//   kNN prior covariances of the block,         lakde_block_posterior (one fused
//   posterior contributions S_n, nu_n, with       kernel for D <= 6; else
//     the lazy rebalance's old column sums,       lakde_knn_prior_cov,
//   factorisation, records, tau update            lakde_posterior_contribs with
//                                                 lakde_rebalance_lazy_begin fused,
//                                                 lakde_posterior)
//   [exact rebalance: old column sums]          lakde_colsum
//   E-step main (+ optional events), with the   lakde_estep_main[_bf16]
//     lazy unsafe-step test                       (+ lakde_rebalance_lazy_check, fused)
//   segment offsets, row pointer, entry count   first half of lakde_estep_finish
//   -- copy of the arena counts and flags --
//   compaction (lazy: storing u, with the       lakde_estep_compact (+ the first half of
//     block's column sums), guarded lazy norm     lakde_rebalance_lazy, fused), lazy_norm
//   -- wait for the copy; host-side outcome: exact fallback / lakde_rebalance --
#include <chrono>

#include "fused_impl.hpp"

using namespace lakde;

namespace {

// Wait for the stream by polling it: the step's wait is on the critical
// path of every partial step, and a blocking synchronisation adds the
// host's wake-up latency to it.
hipError_t wait_stream(hipStream_t s) {
  hipError_t q;
  while ((q = hipStreamQuery(s)) == hipErrorNotReady) {
  }
  return q;
}

hipError_t wait_event(hipEvent_t ev) {
  hipError_t q;
  while ((q = hipEventQuery(ev)) == hipErrorNotReady) {
  }
  return q;
}

// Wait for the flag words of step `seq` (written by the row-pointer kernel
// into host-coherent memory, each word seq << 32 | value) and copy them into
// host_flags.  If the stream completes without the words having arrived,
// the flags are read from device memory instead, so the wait always ends.
// The stream is queried only once the words are late (kFlagPatience): a
// query of a busy stream enqueues a signalled marker, and the device then
// idles ~5 us between the kernels either side of it, every step.
constexpr auto kFlagPatience = std::chrono::milliseconds(20);

int wait_flag_words(const lakde_step_args *a, uint32_t seq, hipStream_t s) {
  volatile unsigned long long *wd = reinterpret_cast<volatile unsigned long long *>(a->flag_words);
  const auto t0 = std::chrono::steady_clock::now();
  bool late = false;
  for (uint64_t it = 0;; ++it) {
    const unsigned long long w0 = wd[0], w1 = wd[1], w2 = wd[2];
    if ((uint32_t)(w0 >> 32) == seq && (uint32_t)(w1 >> 32) == seq &&
        (uint32_t)(w2 >> 32) == seq) {
      a->host_flags[a->block] = (int32_t)(uint32_t)w0;
      a->host_flags[a->num_blocks] = (int32_t)(uint32_t)w1;
      a->host_flags[a->num_blocks + 1] = (int32_t)(uint32_t)w2;
      return LAKDE_OK;
    }
    if ((it & 1023) == 1023) {
      if (!late) {
        late = std::chrono::steady_clock::now() - t0 > kFlagPatience;
        if (!late) continue;
      }
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        const unsigned long long v0 = wd[0], v1 = wd[1], v2 = wd[2];
        if ((uint32_t)(v0 >> 32) == seq && (uint32_t)(v1 >> 32) == seq &&
            (uint32_t)(v2 >> 32) == seq)
          continue;
        LAKDE_HIP(hipMemcpy(a->host_flags, a->nnz_all,
                            (size_t)(a->num_blocks + 2) * sizeof(int32_t), hipMemcpyDeviceToHost));
        return LAKDE_OK;
      }
      if (q != hipErrorNotReady) LAKDE_HIP(q);
    }
  }
}

int step_rebalance(const lakde_step_args *a, hipStream_t s) {
  const lakde_estep_args &e = a->estep;
  const int64_t N = e.N;
  int rc;
  if (a->G == nullptr) {
    return lakde_rebalance(N, a->num_blocks, a->cap, a->block, a->nnz_all, a->cols_all,
                           a->vals_all, e.c_old, e.c, a->rebal_ws, s);
  }
  const int32_t unsafe = a->host_flags[a->num_blocks + 1] & 1;
  if (!unsafe) {
    // the compaction stored u and accumulated the block's sums: normalise
    return rebalance_lazy_guarded(N, nullptr, nullptr, nullptr, e.c_old, e.c, a->G, a->U,
                                  a->U_old, a->lazy_ws, LazyGuard{nullptr, 0, nullptr}, s);
  }
  // the stored form cannot represent this step: rebalance the true values
  // exactly (the other blocks' stored values folded on the way; the step's
  // block holds true values already), recompute the totals, clear the
  // per-step sums
  rc = rebalance_lazy_unsafe(N, a->num_blocks, a->cap, a->block, a->nnz_all, a->cols_all,
                             a->vals_all, e.c_old, e.c, a->G, a->U, a->rebal_ws, s);
  if (rc) return rc;
  LAKDE_HIP(hipMemsetAsync(a->U_old, 0, (size_t)N * sizeof(lakde_fx128), s));
  return LAKDE_OK;
}

// argument checks of one step
int step_check(const lakde_step_args *a) {
  LAKDE_CHECK_ARG(a != nullptr, "null args");
  const lakde_estep_args &e = a->estep;
  const lakde_posterior_args &p = a->post;
  LAKDE_CHECK_ARG(e.mode == LAKDE_ESTEP_UPDATE, "the step runs the UPDATE E-step");
  LAKDE_CHECK_ARG(p.nb == e.nb && p.D == e.D && p.nb >= 1, "posterior and E-step sizes differ");
  LAKDE_CHECK_ARG(a->block >= 0 && a->block < a->num_blocks, "block out of range");
  LAKDE_CHECK_ARG(a->host_flags && a->nnz_all && a->cols_all && a->vals_all, "null pointer");
  LAKDE_CHECK_ARG(e.nnz_dev == a->nnz_all + a->block &&
                      (int32_t *)p.status == a->nnz_all + a->num_blocks,
                  "the block count and status must be slots of nnz_all");
  LAKDE_CHECK_ARG(e.out_col == a->cols_all + a->block * a->cap &&
                      e.out_val == a->vals_all + a->block * a->cap && e.cap == a->cap,
                  "the block's entries must be its arena slots");
  LAKDE_CHECK_ARG(a->G != nullptr || a->rebal_ws != nullptr, "no rebalance workspace");
  LAKDE_CHECK_ARG(a->G == nullptr || (a->U && a->U_old && a->lazy_ws && a->rebal_ws),
                  "lazy rebalance state incomplete");
  return LAKDE_OK;
}

// [fold], prior, contributions (+ the lazy rebalance's old column sums) and
// factorisation [, old column sums, bf16 fragments]; abort: see
// lakde_step_args.abort_flag
int step_begin(const lakde_step_args *a, hipStream_t s, const int32_t *abort) {
  const lakde_estep_args &e = a->estep;
  const lakde_posterior_args &p = a->post;
  const int64_t N = e.N, D = e.D, nb = e.nb;
  int32_t *unsafe = a->nnz_all + a->num_blocks + 1;
  int rc;
  if (a->G != nullptr && a->fold_first) {
    rc = lakde_rebalance_fold(N, a->num_blocks, a->cap, a->nnz_all, a->cols_all, a->vals_all,
                              a->G, a->U, s);
    if (rc) return rc;
  }
  const bool lazy = a->G != nullptr;
  // the block's old column sums are order-independent fixed-point sums
  // (lazy: accumulated by the contributions kernel into the third part of
  // lazy_ws), finished into the floats the E-step reads
  Fx *old_fx = lazy ? a->lazy_ws + 2 * N : nullptr;
  rc = block_posterior(p, e.X, N, e.row0, e.rowptr, e.out_col, e.out_val, a->G, a->knn_graph,
                       a->knn_graph_dtype, a->knn_k, a->nu0, old_fx, lazy ? a->U_old : nullptr,
                       lazy ? unsafe : nullptr, s, abort);
  if (rc) return rc;
  if (lazy) {
    rc = colsum_finish(N, old_fx, const_cast<float *>(e.old_colsum), s);
  } else {
    Fx *acc = (Fx *)((char *)a->rebal_ws + ((size_t)N * 4 + 15) / 16 * 16);
    rc = lakde_colsum(e.out_col, e.out_val, e.nnz_dev, N, const_cast<float *>(e.old_colsum), acc,
                      s);
  }
  if (rc) return rc;
  if (a->bf16_frag != nullptr) {
    rc = a->fp32_mfma            ? lakde_pack_m32(e.rec, nb, D, a->screen_origin, a->bf16_frag, s)
         : a->screen_xb != nullptr ? lakde_pack_screen(e.rec, nb, D, a->screen_origin,
                                                       a->bf16_frag, s)
                                   : lakde_pack_frag_bf16(e.rec, nb, D, a->bf16_frag, s);
    if (rc) return rc;
  }
  return LAKDE_OK;
}

// E-step, offsets (+ flags), compaction and the guarded lazy normalisation;
// *seq: the flag words' tag of this step (0 without flag words)
int step_rest(const lakde_step_args *a, hipStream_t s, uint32_t *seq_out) {
  const lakde_estep_args &e = a->estep;
  const int64_t N = e.N;
  int32_t *unsafe = a->nnz_all + a->num_blocks + 1;
  int rc;
  // the E-step main kernel also runs the lazy rebalance's unsafe-step test
  lakde_estep_args em = e;
  if (a->G != nullptr) {
    em.lazy_G = a->G;
    em.lazy_unsafe = unsafe;
    if (a->rebal_ws != nullptr) em.lazy_factor = rebase_factors(N, a->rebal_ws);
  }
  if (a->ev_main[0]) LAKDE_HIP(hipEventRecord((hipEvent_t)a->ev_main[0], s));
  if (a->bf16_frag != nullptr && a->fp32_mfma)
    rc = lakde_estep_main_m32(&em, a->bf16_frag, a->screen_xnorm, s);
  else if (a->bf16_frag != nullptr && a->screen_xb != nullptr && a->screen_split_ws != nullptr)
    rc = lakde_estep_main_screen_split(&em, a->bf16_frag, a->screen_xb, a->screen_xnorm,
                                       a->screen_split_ws, a->screen_split_bytes, s);
  else if (a->bf16_frag != nullptr && a->screen_xb != nullptr)
    rc = lakde_estep_main_screen(&em, a->bf16_frag, a->screen_xb, a->screen_xnorm, nullptr, s);
  else if (a->bf16_frag != nullptr)
    rc = lakde_estep_main_bf16(&em, a->bf16_frag, s);
  else
    rc = lakde_estep_main(&em, s);
  if (rc) return rc;
  if (a->ev_main[1]) LAKDE_HIP(hipEventRecord((hipEvent_t)a->ev_main[1], s));
  // offsets first: the flags the host needs are final before the compaction;
  // with flag_words the row-pointer kernel hands them to the host itself,
  // otherwise one copy (and optionally an event to wait on)
  uint32_t seq = 0;
  if (a->flag_words != nullptr) {
    volatile unsigned long long *wd =
        reinterpret_cast<volatile unsigned long long *>(a->flag_words);
    seq = (uint32_t)wd[3] + 1u;
    if (seq == 0u) seq = 1u;
    wd[3] = seq;
    // the kernel writes through the words' device-side address (word 4)
    auto *dev_words = reinterpret_cast<unsigned long long *>((uintptr_t)wd[4]);
    rc = estep_count(e, s, dev_words, a->nnz_all + a->num_blocks, unsafe, seq, a->abort_flag);
    if (rc) return rc;
  } else {
    rc = estep_count(e, s);
    if (rc) return rc;
    LAKDE_HIP(hipMemcpyAsync(a->host_flags, a->nnz_all,
                             (size_t)(a->num_blocks + 2) * sizeof(int32_t),
                             hipMemcpyDeviceToHost, s));
    if (a->ev_flags) LAKDE_HIP(hipEventRecord((hipEvent_t)a->ev_flags, s));
  }
  // the compaction (skips itself when the entries do not fit) and the usual
  // rebalance outcome run while the host reads the flags and prepares the
  // next step; the lazy kernels skip themselves for an unsafe or
  // overflowing step
  if (a->G != nullptr && a->rebal_ws != nullptr) {
    // columns to re-base (nothing unless only bit 1 of the flag is set)
    rc = rebalance_lazy_rebase(N, a->num_blocks, a->cap, a->block, a->nnz_all, a->cols_all,
                               a->vals_all, a->G, a->U, a->U_old, a->rebal_ws, unsafe, s);
    if (rc) return rc;
  }
  if (a->G != nullptr) {
    // the compaction also stores u = r / (G coeff) and the block's column
    // sums unless the step is unsafe; the normalisation skips itself for an
    // unsafe or overflowing step
    rc = estep_compact_lazy(e, a->G, unsafe, a->lazy_ws, a->lazy_ws + N, s);
    if (rc) return rc;
    rc = rebalance_lazy_guarded(N, nullptr, nullptr, nullptr, e.c_old, e.c, a->G, a->U,
                                a->U_old, a->lazy_ws, LazyGuard{e.nnz_dev, a->cap, unsafe}, s);
    if (rc) return rc;
  } else {
    rc = lakde_estep_compact(&e, s);
    if (rc) return rc;
  }
  *seq_out = seq;
  return LAKDE_OK;
}

// wait for the step's flags; the host's share of the step
int step_wait(const lakde_step_args *a, hipStream_t s, uint32_t seq) {
  int rc;
  if (a->flag_words != nullptr) {
    rc = wait_flag_words(a, seq, s);
    if (rc) return rc;
  } else {
    LAKDE_HIP(a->ev_flags ? wait_event((hipEvent_t)a->ev_flags) : wait_stream(s));
  }
  if (a->host_flags[a->num_blocks] != 0) {
    set_error("lakde_partial_step: Stabilization did not converge");
    return LAKDE_ESTABILIZE;
  }
  if (a->host_flags[a->block] > a->cap) {
    set_error("lakde_partial_step: the block's entries exceed the arena capacity");
    return LAKDE_ECAPACITY;
  }
  if (a->G != nullptr && (a->host_flags[a->num_blocks + 1] & 1) == 0) return LAKDE_OK;  // done
  return step_rebalance(a, s);
}

// whether the next step's posterior may be enqueued before this step's
// flags (the fused D <= 6 posterior honours the abort word)
bool can_begin_ahead(const lakde_step_args *a) {
  return a->G != nullptr && a->bf16_frag == nullptr && a->flag_words != nullptr &&
         a->abort_flag != nullptr && a->estep.D <= 6;
}

}  // namespace

extern "C" {

int lakde_partial_step(const lakde_step_args *a, lakde_stream_t stream) {
  int rc = step_check(a);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  rc = step_begin(a, s, nullptr);
  if (rc) return rc;
  uint32_t seq = 0;
  rc = step_rest(a, s, &seq);
  if (rc) return rc;
  return step_wait(a, s, seq);
}

int lakde_partial_sweep(lakde_step_args *steps, int64_t first, int64_t count, float *const *c,
                        int *cur, int *lazy_dirty, int64_t *done, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(steps && c && c[0] && c[1] && cur && lazy_dirty && done, "null pointer");
  LAKDE_CHECK_ARG(*cur == 0 || *cur == 1, "cur must be 0 or 1");
  LAKDE_CHECK_ARG(first >= 0 && count >= 0, "bad range");
  hipStream_t s = as_stream(stream);
  *done = 0;
  bool begun = false;  // the current step's posterior was enqueued ahead, guarded
  for (int64_t i = 0; i < count; ++i) {
    lakde_step_args *a = steps + first + i;
    LAKDE_CHECK_ARG(a->block == first + i, "steps[b] must be the arguments of block b");
    a->estep.c_old = c[*cur];
    a->estep.c = c[1 - *cur];
    a->fold_first = (a->G != nullptr && a->block == 0 && *lazy_dirty) ? 1 : 0;
    int rc = step_check(a);
    if (rc) return rc;
    if (!begun) {
      rc = step_begin(a, s, nullptr);
      if (rc) return rc;
    }
    uint32_t seq = 0;
    rc = step_rest(a, s, &seq);
    if (rc) return rc;
    // the next step's posterior goes in before the wait, so the device does
    // not idle while the host reads this step's flags; it does nothing if
    // this step turns out to need the host (the abort word)
    begun = false;
    if (i + 1 < count && can_begin_ahead(a)) {
      lakde_step_args *nx = a + 1;
      LAKDE_CHECK_ARG(nx->block == first + i + 1, "steps[b] must be the arguments of block b");
      nx->estep.c_old = c[1 - *cur];
      nx->estep.c = c[*cur];
      nx->fold_first = 0;
      if (can_begin_ahead(nx) && nx->abort_flag == a->abort_flag) {
        rc = step_check(nx);
        if (rc) return rc;
        rc = step_begin(nx, s, a->abort_flag);
        if (rc) return rc;
        begun = true;
      }
    }
    rc = step_wait(a, s, seq);
    if (rc) return rc;
    // an unsafe step was finished by the host after the guarded posterior,
    // which did nothing: enqueue it again
    if (a->G != nullptr && (a->host_flags[a->num_blocks + 1] & 1) != 0) begun = false;
    *cur = 1 - *cur;
    if (a->G != nullptr) *lazy_dirty = (a->host_flags[a->num_blocks + 1] & 1) == 0 ? 1 : 0;
    *done = i + 1;
  }
  return LAKDE_OK;
}

int lakde_partial_step_resume(const lakde_step_args *a, lakde_stream_t stream) {
  LAKDE_CHECK_ARG(a != nullptr, "null args");
  LAKDE_CHECK_ARG(a->host_flags[a->block] <= a->cap, "capacity still too small");
  LAKDE_CHECK_ARG(a->estep.out_col == a->cols_all + a->block * a->cap &&
                      a->estep.out_val == a->vals_all + a->block * a->cap &&
                      a->estep.cap == a->cap,
                  "the block's entries must be its arena slots");
  hipStream_t s = as_stream(stream);
  int rc = a->G != nullptr
               ? estep_compact_lazy(a->estep, a->G, a->nnz_all + a->num_blocks + 1, a->lazy_ws,
                                    a->lazy_ws + a->estep.N, s)
               : lakde_estep_compact(&a->estep, s);
  if (rc) return rc;
  return step_rebalance(a, s);
}

}  // extern "C"
