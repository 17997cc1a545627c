"""Locally adaptive KDE core on the fused HIP path
(mirrors lakde/local_adaptive_kde.py:47-314).

Per partial step of kernel block b the device work is:

  prior (kNN covariance or the shared prior)           lakde_knn_prior_cov
  posterior contributions S_n, nu_n                    lakde_posterior_contribs
  factorisation, L^{-1}, logdet, trace, tau update     lakde_posterior
  old column sums of block b                           lakde_rebalance_lazy_begin
  quad forms + LSE + renormalise + compaction to CSR   lakde_estep
  rebalance                                            lakde_rebalance_lazy

with one host synchronisation (the new block's entry count and the
stabilisation status).  Responsibilities live in an ``Arena``.

The rebalance (lakde/local_adaptive_kde.py:156-181) rescales every column
of every block by per-column factors.  The default ``rebalance="lazy"``
keeps those factors as a per-column vector G (true value = stored value *
G[col]) so that a partial step touches only the updated block's entries,
and folds G into the values once per sweep and whenever the values are
handed out; ``rebalance="exact"`` rescales the whole arena every step
(lakde_colsum + lakde_rebalance), as the reference does.
"""

import math

import torch

from . import _fused as F
from ._lib import ESTEP_ACCUM, ESTEP_FIXED, ESTEP_UPDATE, LakdeError, require_gpu
from .kdes import AbstractKDE
from .utils import tensor_as


class LocallyAdaptiveKDE(AbstractKDE):
    """``precision``: "fp32" (VALU quadratic forms) or "bf16" (the precision
    form: quadratic forms on the BF16 matrix cores with bf16x3 split
    precision, kept responsibilities recomputed exactly in FP32; 8 <= D <= 63).
    """

    def __init__(self, nu_0, a_0, b_0, block_size, verbose, logs, precision="fp32",
                 rebalance="lazy"):
        super().__init__(block_size, verbose, logs)
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        if rebalance not in ("lazy", "exact"):
            raise ValueError("rebalance must be 'lazy' or 'exact'")
        self.precision = precision
        self.rebalance = rebalance
        self._lazy = None
        self._unsafe_step = False
        self.nu_0 = nu_0
        self.a_0 = a_0
        self.b_0 = b_0
        self.nu = None
        self.a = None
        self.b = None
        self.arena = None
        self.rnm_log_consts = None
        self.logdet_W = None
        self.W_inv_jitter = None
        self.rnm_active_contribs = None
        self._estep_timer = None  # optional callable wrapping the E-step main launch

    # ------------------------------------------------------------ helpers
    def _blocks(self, N):
        return range((N - 1) // self.block_size + 1)

    def _span(self, b, N):
        s = b * self.block_size
        e = min(s + self.block_size, N)
        return s, e, e - s

    def _common_init(self, X, sparsity_threshold):
        require_gpu(X)
        if X.dtype != torch.float32:
            raise TypeError("lakde_amd fits float32 data")
        N, D = X.shape
        if D > 63:
            raise ValueError("the fused path supports D <= 63")
        self.block_size = min(self.block_size, N)
        self.W_inv_jitter = X.new_zeros(N)
        self.a_0 = tensor_as(self.a_0, X)
        self.b_0 = tensor_as(self.b_0, X)
        if self.nu_0 is not None:
            self.nu_0 = tensor_as(self.nu_0, X)
        else:
            self.nu_0 = X.new_full((), D)
        if torch.any(self.nu_0 <= D - 1):
            raise ValueError("nu_0 must be greater than the data dimensionality minus one, "
                             "or the ELBO will be undefined")
        self._nu0 = float(self.nu_0)
        self._b0 = float(self.b_0)

    def _alloc(self, X):
        N, D = X.shape
        bs = self.block_size
        dev = X.device
        self._R = F.record_floats(D)
        self.rec_D = D
        self.rec = torch.zeros(N, self._R, device=dev)
        self._S = torch.empty(bs, D, D, device=dev)
        self._prior_buf = torch.empty(bs, D, D, device=dev)
        self._prior_tril = torch.empty(bs, D, D, device=dev)
        self._Linv = torch.empty(bs, D, D, device=dev)
        self._nu_s = torch.empty(bs, device=dev)
        self._logdet_s = torch.empty(bs, device=dev)
        self._logdet_prior_s = torch.empty(bs, device=dev)
        self._trace_s = torch.empty(bs, device=dev)
        self._c_new = torch.empty(N, device=dev)
        if self.rebalance == "lazy":
            self._lazy = F.LazyRebalance(N, dev)
        else:
            self._colsum = torch.empty(N, device=dev)
            self._rebal_ws = torch.empty(2 * N, device=dev)
        self._ews = F.EstepWorkspace(N, bs, dev)
        if self.precision == "bf16":
            if not 8 <= D <= 63:
                raise ValueError("the bf16 precision form needs 8 <= D <= 63")
            self._fb = F.bf16_frag_bytes(D)
            self._frag_blk = torch.empty(bs * self._fb, dtype=torch.uint8, device=dev)

    def _init_arena(self, X, graph, sparsity_threshold):
        """Initial responsibilities 1/init_k on the kNN graph, laid out per
        kernel block in canonical order (lakde/sparse_util.py:8-54,
        lakde/local_knn_kde.py:113-127)."""
        N = X.size(0)
        bs = self.block_size
        if sparsity_threshold:
            init_k = min(math.floor(1 / sparsity_threshold), graph.size(-1))
            g = graph[:, :init_k]
        else:
            g, init_k = graph, graph.size(-1)
        K = g.size(1)
        row = g.reshape(-1).long()
        col = torch.arange(N, device=X.device).repeat_interleave(K)
        order = torch.argsort(row * N + col)
        row, col = row[order], col[order]
        blk = row // bs
        nblocks = (N - 1) // bs + 1
        counts = torch.bincount(blk, minlength=nblocks)
        starts = torch.cumsum(counts, 0) - counts
        pos = torch.arange(row.numel(), device=X.device) - starts[blk]
        cap = int(counts.max()) if row.numel() else 1
        arena = F.Arena(nblocks, bs, max(256, (cap * 3 // 2 + 255) // 256 * 256), X.device)
        arena.row[blk, pos] = (row - blk * bs).to(torch.int32)
        arena.col[blk, pos] = col.to(torch.int32)
        arena.val[blk, pos] = float(torch.tensor(1.0 / init_k, dtype=torch.float32))
        arena.nnz = [int(c) for c in counts.tolist()]
        arena.nnz_dev[:nblocks] = counts.to(torch.int32)
        from .kernels import ind2ptr
        for b in range(nblocks):
            s, e, nb = self._span(b, N)
            ind2ptr(arena.row[b, :arena.nnz[b]], nb, out=arena.rowptr[b, :nb + 1])
        self.arena = arena
        self.rnm_active_contribs = X.new_tensor(float(sum(arena.nnz)))

    # ---------------------------------------------------- posterior of a block
    def _prior(self, X, b, nb):
        """Returns (prior_mode, prior, prior_jitter, prior_tril, logdet_prior)."""
        raise NotImplementedError

    def _block_posterior(self, X, b, nu_out, logdet_out, trace_out, logdet_prior_out,
                         update_b):
        """Posterior of block b (lakde/local_adaptive_kde.py:66-96): writes the
        records rec[s:e]; optionally the tau update of b[s:e]."""
        N, D = X.shape
        s, e, nb = self._span(b, N)
        arena = self.arena
        S = self._S[:nb]
        F.posterior_contribs(arena.rowptr[b, :nb + 1], arena.col[b], arena.val[b], X[s:e], X,
                             self._nu0, S, nu_out, colscale=self._colscale())
        mode, prior, prior_jitter, prior_tril = self._prior(X, b, nb)
        F.posterior(mode, prior, prior_jitter, S, nu_out, self.a[s:e], self.b[s:e], self._nu0,
                    self.W_inv_jitter[s:e], X[s:e], self._Linv[:nb], prior_tril, logdet_out,
                    logdet_prior_out, trace_out, self.rec[s:e], arena.status_ptr,
                    b0=self._b0, b_out=self.b[s:e] if update_b else None)

    def _sync(self, b):
        flags = self.arena.nnz_dev.cpu()
        nblk = self.arena.num_blocks
        if int(flags[nblk]) != 0:
            raise RuntimeError("Stabilization did not converge")
        self._unsafe_step = int(flags[nblk + 1]) != 0
        return int(flags[b]) if b is not None else None

    def _estep_into_arena(self, X, b, mode, threshold, c_old=None, old_colsum=None):
        N = X.size(0)
        s, e, nb = self._span(b, N)
        c = self._c_new if mode == ESTEP_UPDATE else self.rnm_log_consts
        F.estep(X, self.rec[s:e], s, mode, threshold, c, self._ews, c_old=c_old,
                old_colsum=old_colsum, arena=self.arena, b=b, timer=self._estep_timer,
                bf16=self._bf16_operands(b, N))
        if mode == ESTEP_UPDATE and self._lazy is not None:
            self._lazy.check(self.arena, c_old, c)
        nnz = self._sync(b)
        if nnz > self.arena.cap:
            self.arena.grow(nnz)
            F.estep(X, self.rec[s:e], s, mode, threshold, c, self._ews, c_old=c_old,
                    old_colsum=old_colsum, arena=self.arena, b=b, compact_only=True)
        old = self.arena.nnz[b]
        self.arena.nnz[b] = nnz
        return old, nnz

    def _bf16_operands(self, b, N):
        if self.precision != "bf16":
            return None
        s, e, nb = self._span(b, N)
        frag = self._frag_blk[:nb * self._fb]
        F.pack_frag_bf16(self.rec[s:e], self.rec_D, frag)
        return frag

    # ------------------------------------------------------------- E-step
    def _init_lse_pass(self, X, b):
        """lakde/local_knn_kde.py:157-166: c = logaddexp(c, LSE over block b)."""
        N = X.size(0)
        s, e, nb = self._span(b, N)
        F.estep(X, self.rec[s:e], s, ESTEP_ACCUM, 0.0, self.rnm_log_consts, self._ews,
                bf16=self._bf16_operands(b, N))

    def _init_fixed_pass(self, X, b, sparsity_threshold):
        """partial_expectation_step(update_normalization=False)."""
        old, new = self._estep_into_arena(X, b, ESTEP_FIXED, sparsity_threshold)
        self.rnm_active_contribs += new

    def _update_step(self, X, b, sparsity_threshold):
        """partial_expectation_step + rebalance_responsibilities_
        (lakde/local_adaptive_kde.py:98-181)."""
        N = X.size(0)
        arena = self.arena
        lazy = self._lazy
        if lazy is not None:
            if b == 0 and lazy.dirty:
                lazy.fold(arena)  # once per sweep: bounds the drift of G and U
            lazy.begin(arena, b)
            old_colsum = lazy.old_colsum
        else:
            F.colsum(arena.col[b], arena.val[b], arena.nnz_dev[b:b + 1], N, self._colsum)
            old_colsum = self._colsum
        old, new = self._estep_into_arena(X, b, ESTEP_UPDATE, sparsity_threshold,
                                          c_old=self.rnm_log_consts, old_colsum=old_colsum)
        self.rnm_active_contribs += new - old
        if lazy is not None:
            if self._unsafe_step:
                lazy.exact_step(arena, b, self.rnm_log_consts, self._c_new)
                lazy.exact_steps += 1
            else:
                lazy.end(arena, b, self.rnm_log_consts, self._c_new)
        else:
            F.rebalance(arena, b, self.rnm_log_consts, self._c_new, self._rebal_ws)
        self.rnm_log_consts, self._c_new = self._c_new, self.rnm_log_consts

    def _init_rebalance(self):
        if self._lazy is not None:
            self._lazy.fold(self.arena)  # G = 1, U = exact column sums
            self._lazy.end(self.arena, None, self.rnm_log_consts, self.rnm_log_consts)
        else:
            F.rebalance(self.arena, -1, self.rnm_log_consts, self.rnm_log_consts,
                        self._rebal_ws)

    def _colscale(self):
        """Per-column factors of the stored values (None when they are 1)."""
        return self._lazy.G if self._lazy is not None and self._lazy.dirty else None

    def materialize(self):
        """Fold pending lazy-rebalance factors into the stored values, so
        that the arena holds the true responsibilities."""
        if self._lazy is not None and self._lazy.dirty:
            self._lazy.fold(self.arena)

    def fit(self, X, iterations=200, callbacks=None, sparsity_threshold=None, dataset_label=None):
        super().fit(X, iterations, callbacks, sparsity_threshold, dataset_label)
        self.materialize()
        return self

    # ---------------------------------------------------------- prediction
    def log_pred_density(self, X, Y):
        """log p(y) = -ln N + LSE_n log St(y | posterior kernel n)
        (lakde/local_adaptive_kde.py:200-223).  Recomputes every block's
        posterior (updating the persistent jitters, as the reference does),
        then streams all N kernels over the test points in one launch."""
        require_gpu(X, Y)
        if self.arena is None:
            raise RuntimeError("fit the model first")
        N, D = X.shape
        Y = Y.contiguous()
        if Y.dtype != torch.float32:
            raise TypeError("Y must be float32")
        bs = self.block_size
        for b in self._blocks(N):
            s, e, nb = self._span(b, N)
            self._block_posterior(X, b, self._nu_s[:nb], self._logdet_s[:nb],
                                  self._trace_s[:nb], self._logdet_prior_s[:nb], update_b=False)
        self._sync(None)
        out = torch.empty(Y.size(0), device=Y.device)
        if self.precision == "bf16":
            frag = torch.empty(N * self._fb, dtype=torch.uint8, device=Y.device)
            F.pack_frag_bf16(self.rec, D, frag)
            return F.pred_density_bf16(self.rec, frag, Y, out)
        return F.pred_density(self.rec, Y, out)

    # ------------------------------------------------------------- state
    @property
    def rnm_blocks(self):
        """Per-block (indices (2, nnz) int32, values) copies, the reference
        layout (rows local to the block)."""
        self.materialize()
        return [tuple(t.clone() for t in self.arena.block(b))
                for b in range(self.arena.num_blocks)]

    def data_log_likelihood_no_rnm(self, X):
        N, D = X.shape
        return -D / 2 * math.log(2 * math.pi) + self.rnm_log_consts

    def state_dict(self):
        d = super().state_dict()
        d.update({
            "nu_0": self.nu_0, "nu": self.nu, "a_0": self.a_0, "b_0": self.b_0,
            "a": self.a, "b": self.b, "rnm_blocks": self.rnm_blocks,
            "rnm_log_consts": self.rnm_log_consts, "logdet_W": self.logdet_W,
            "rnm_active_contribs": self.rnm_active_contribs, "W_inv_jitter": self.W_inv_jitter,
        })
        return d

    def _load_common(self, state_dict, X_device):
        super().load_state_dict(state_dict)
        dev = X_device
        for k in ("nu_0", "nu", "a_0", "b_0", "a", "b", "rnm_log_consts", "logdet_W",
                  "rnm_active_contribs", "W_inv_jitter"):
            v = state_dict[k]
            setattr(self, k, v.to(dev).clone() if torch.is_tensor(v) else v)
        self.a = self.a.contiguous()
        self._nu0 = float(self.nu_0)
        self._b0 = float(self.b_0)

    def _load_blocks(self, blocks, X):
        """Rebuild the arena from reference-layout (indices, values) blocks,
        sorting each block into canonical order."""
        N = X.size(0)
        nblocks = len(blocks)
        cap = max(256, max(int(v.numel()) for _, v in blocks))
        arena = F.Arena(nblocks, self.block_size, (cap * 5 // 4 + 255) // 256 * 256, X.device)
        from .kernels import ind2ptr
        for b, (ind, val) in enumerate(blocks):
            ind = ind.to(X.device)
            val = val.to(X.device, torch.float32)
            order = torch.argsort(ind[0].long() * N + ind[1].long())
            n = int(val.numel())
            arena.row[b, :n] = ind[0][order].to(torch.int32)
            arena.col[b, :n] = ind[1][order].to(torch.int32)
            arena.val[b, :n] = val[order]
            arena.nnz[b] = n
            s, e, nb = self._span(b, N)
            ind2ptr(arena.row[b, :n], nb, out=arena.rowptr[b, :nb + 1])
        arena.nnz_dev[:nblocks] = torch.tensor(arena.nnz, dtype=torch.int32, device=X.device)
        self.arena = arena

    def _finish_load(self, X):
        self._alloc(X)
        if self._lazy is not None:
            self._lazy.fold(self.arena)  # G = 1, U = exact column sums
