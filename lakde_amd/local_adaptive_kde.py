"""Locally adaptive KDE core on the fused HIP path
(mirrors lakde/local_adaptive_kde.py:47-314).

A partial step of kernel block b is one call of the native step driver
(lakde_partial_step), which launches

  prior (kNN covariance; the shared prior is an input) lakde_knn_prior_cov
  posterior contributions S_n, nu_n                    lakde_posterior_contribs
  factorisation, L^{-1}, logdet, trace, tau update     lakde_posterior
  old column sums of block b                           lakde_rebalance_lazy_begin
  quad forms + LSE + renormalise + compaction to CSR   lakde_estep
  rebalance                                            lakde_rebalance_lazy

with one host synchronisation (the new block's entry count, the
stabilisation status, the lazy rebalance's unsafe-step flag).
Responsibilities live in an ``Arena``.

The rebalance (lakde/local_adaptive_kde.py:156-181) rescales every column
of every block by per-column factors.  The default ``rebalance="lazy"``
keeps those factors as a per-column vector G (true value = stored value *
G[col]) so that a partial step touches only the updated block's entries,
and folds G into the values once per sweep and whenever the values are
handed out; ``rebalance="exact"`` rescales the whole arena every step
(lakde_colsum + lakde_rebalance), as the reference does.
"""

import ctypes
import math

import torch

from . import _fused as F
from . import _lib
from ._lib import (ESTEP_ACCUM, ESTEP_FIXED, ESTEP_UPDATE, LAKDE_ECAPACITY, LAKDE_ESTABILIZE,
                   LAKDE_OK, LakdeError, check, require_gpu, stream_of)
from .kdes import AbstractKDE
from .utils import InterruptDelay, column_order, tensor_as


class LocallyAdaptiveKDE(AbstractKDE):
    """``precision``: "fp32" (FP32 quadratic forms for every pair) or "bf16"
    (the precision form, 8 <= D <= 63: the E-step screens every pair on the
    BF16 matrix cores with a rigorous error bound and recomputes exactly in
    FP32 the pairs that can matter to a normaliser or a kept responsibility;
    the predictive density runs bf16x3 split precision on the matrix cores,
    or, with ``pred_screen``, screens the same way as the E-step and
    evaluates the terms that can matter exactly in FP32).
    The FP32 E-step runs on the FP32 matrix cores for 17 <= D <= 63
    (``fp32_mfma``, on by default), on the vector ALUs otherwise.
    """

    fp32_mfma = True
    # precision form: the predictive density with bf16x3 terms for every
    # pair (False: 1.4e-4 nats from float64 at mix128_d63, 31k test points/s)
    # or with the BF16 screen and exact FP32 terms for the pairs within the
    # margin (True: 3.8e-6 nats, 18k test points/s; the Student-t terms fall
    # off only logarithmically, so about a tenth of the pairs need the
    # exact term)
    pred_screen = False
    # precision form: the E-step's UPDATE launches in the split form (the
    # block's kernels in ranges screened by separate workgroups, combined
    # per column)
    screen_split = True
    _screen = None

    def __init__(self, nu_0, a_0, b_0, block_size, verbose, logs, precision="fp32",
                 rebalance="lazy", estep_skip=True):
        super().__init__(block_size, verbose, logs)
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        if rebalance not in ("lazy", "exact"):
            raise ValueError("rebalance must be 'lazy' or 'exact'")
        self.precision = precision
        self.rebalance = rebalance
        self.estep_skip = bool(estep_skip)
        self._col_order = None
        self._lazy = None
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
        self._active = None         # rnm_active_contribs, kept on the host
        self._active_device = None
        self._step_cache = {}
        self._estep_timer = None  # optional: .pair() -> two recorded torch events or None
        self._count_steps = False
        # partial steps per native sweep call in fit(): bounds the delay of
        # a SIGINT (it takes effect between calls)
        self.sweep_chunk = 256

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
            self._rebal_ws = F.rebalance_workspace(N, dev)
        self._host_flags = torch.zeros(self.arena.num_blocks + 2, dtype=torch.int32,
                                       pin_memory=True)
        self._flags = self._host_flags.numpy()
        self._flag_event = torch.cuda.Event()  # the driver waits on it for the flags copy
        self._flag_event.record()
        self._flag_words = F.FlagWords()  # ... or, when given, on these flag words
        self._step_cache = {}
        self._ews = F.EstepWorkspace(N, bs, dev)
        # pair skipping (FP32, D <= 16): spatially compact column chunks
        self._col_order = None
        if self.estep_skip and self.precision == "fp32" and D <= 16:
            self._col_order = column_order(X)
        self._screen = None
        if self.precision == "bf16" and not 8 <= D <= 63:
            raise ValueError("the bf16 precision form needs 8 <= D <= 63")
        if 8 <= D <= 63:
            # the matrix-core E-steps (the screening form; FP32 for D >= 17):
            # per-fit points, per-block fragments (either form may run: the
            # precision can change between steps)
            self._screen = F.ScreenPoints(X)
            self._fb = F.screen_frag_bytes(D)
            self._m32_fb = F.m32_frag_bytes(D) if D >= 17 else 0
            self._frag_blk = torch.empty(bs * max(self._fb, self._m32_fb), dtype=torch.uint8,
                                         device=dev)
            # the screening E-step's split form (UPDATE steps): the block's
            # kernels in ranges screened by separate workgroups
            self._split_bytes = int(_lib.load().lakde_screen_split_workspace(N, bs))
            self._split_ws = torch.empty(max(self._split_bytes, 1), dtype=torch.uint8,
                                         device=dev)

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
        self._step_cache = {}
        self._active_device = X.device
        self._active = float(sum(arena.nnz))

    @property
    def rnm_active_contribs(self):
        """Number of kept responsibilities (lakde/local_adaptive_kde.py:146)."""
        if self._active is None:
            return None
        return torch.tensor(self._active, dtype=torch.float32, device=self._active_device)

    @rnm_active_contribs.setter
    def rnm_active_contribs(self, v):
        self._active = None if v is None else float(v)
        if torch.is_tensor(v):
            self._active_device = v.device

    # ---------------------------------------------------- posterior of a block
    def _prior_spec(self, X, b, nb):
        """(prior_mode, prior, prior_jitter, prior_tril, knn_graph rows or None)."""
        raise NotImplementedError

    def _step_outputs(self, b, nb):
        """Posterior outputs of a partial step: (nu, logdet_W, trace, logdet_prior)."""
        raise NotImplementedError

    def _ensure_prior_cache(self, X):
        """Fill a per-fit prior cache before the first posterior (kNN model)."""

    def _block_posterior(self, X, b, nu_out, logdet_out, trace_out, logdet_prior_out,
                         update_b):
        """Posterior of block b (lakde/local_adaptive_kde.py:66-96): writes the
        records rec[s:e]; optionally the tau update of b[s:e]."""
        self._ensure_prior_cache(X)
        N, D = X.shape
        s, e, nb = self._span(b, N)
        arena = self.arena
        mode, prior, prior_jitter, prior_tril, graph = self._prior_spec(X, b, nb)
        post = F.posterior_args(mode, prior, prior_jitter, self._S[:nb], nu_out, self.a[s:e],
                                self.b[s:e], self._nu0, self.W_inv_jitter[s:e], X[s:e],
                                self._Linv[:nb], prior_tril, logdet_out, logdet_prior_out,
                                trace_out, self.rec[s:e], arena.status_ptr, b0=self._b0,
                                b_out=self.b[s:e] if update_b else None)
        F.block_posterior(X, s, arena.rowptr[b, :nb + 1], arena.col[b], arena.val[b],
                          self._colscale(), graph, post)

    def _sync(self, b):
        flags = self.arena.nnz_dev.cpu()
        if int(flags[self.arena.num_blocks]) != 0:
            raise RuntimeError("Stabilization did not converge")
        return int(flags[b]) if b is not None else None

    def _estep_into_arena(self, X, b, mode, threshold):
        """A FIXED-mode E-step of block b into the arena (initialisation)."""
        N = X.size(0)
        s, e, nb = self._span(b, N)
        c = self.rnm_log_consts
        F.estep(X, self.rec[s:e], s, mode, threshold, c, self._ews, arena=self.arena, b=b,
                screen=self._screen_operands(b, N), m32=self._m32_operands(b, N),
                order=self._col_order)
        nnz = self._sync(b)
        if nnz > self.arena.cap:
            self.arena.grow(nnz)
            self._step_cache = {}
            F.estep(X, self.rec[s:e], s, mode, threshold, c, self._ews, arena=self.arena, b=b,
                    compact_only=True, order=self._col_order)
        old = self.arena.nnz[b]
        self.arena.nnz[b] = nnz
        return old, nnz

    def _screen_operands(self, b, N):
        """(fragments of block b, the fit's screening points) for the
        screening E-step, or None for the FP32 path."""
        if self.precision != "bf16":
            return None
        s, e, nb = self._span(b, N)
        frag = self._frag_blk[:nb * self._fb]
        F.pack_screen(self.rec[s:e], self.rec_D, self._screen.origin, frag)
        return frag, self._screen

    def _uses_m32(self):
        return (self.precision == "fp32" and self.fp32_mfma and self._screen is not None
                and self.rec_D >= 17 and self._col_order is None)

    def _m32_operands(self, b, N):
        """(fragments of block b, the fit's points) for the FP32 matrix-core
        E-step, or None."""
        if not self._uses_m32():
            return None
        s, e, nb = self._span(b, N)
        frag = self._frag_blk[:nb * self._m32_fb]
        F.pack_m32(self.rec[s:e], self.rec_D, self._screen.origin, frag)
        return frag, self._screen

    # ------------------------------------------------------------- E-step
    def _init_lse_pass(self, X, b):
        """lakde/local_knn_kde.py:157-166: c = logaddexp(c, LSE over block b)."""
        N = X.size(0)
        s, e, nb = self._span(b, N)
        F.estep(X, self.rec[s:e], s, ESTEP_ACCUM, 0.0, self.rnm_log_consts, self._ews,
                screen=self._screen_operands(b, N), m32=self._m32_operands(b, N))

    def _init_fixed_pass(self, X, b, sparsity_threshold):
        """partial_expectation_step(update_normalization=False)."""
        old, new = self._estep_into_arena(X, b, ESTEP_FIXED, sparsity_threshold)
        self._active += new

    # --------------------------------------------------------- partial step
    def _step_args(self, X, b):
        """The step driver's arguments of block b, an element of one array
        over all blocks built once per arena layout; the per-step fields are
        set by _run_steps and the driver."""
        arena = self.arena
        arr = self._step_cache.get("arr")
        if arr is None:
            arr = (_lib.StepArgs * arena.num_blocks)()
            self._step_cache = {"arr": arr, "built": [False] * arena.num_blocks}
        if self._step_cache["built"][b]:
            return arr[b]
        N, D = X.shape
        s, e, nb = self._span(b, N)
        mode, prior, prior_jitter, prior_tril, graph = self._prior_spec(X, b, nb)
        nu_out, logdet_out, trace_out, logdet_prior_out = self._step_outputs(b, nb)
        lazy = self._lazy
        a = arr[b]
        a.post = F.posterior_args(mode, prior, prior_jitter, self._S[:nb], nu_out, self.a[s:e],
                                  self.b[s:e], self._nu0, self.W_inv_jitter[s:e], X[s:e],
                                  self._Linv[:nb], prior_tril, logdet_out, logdet_prior_out,
                                  trace_out, self.rec[s:e], arena.status_ptr, b0=self._b0,
                                  b_out=self.b[s:e])
        a.estep = F.estep_args(X, self.rec[s:e], s, ESTEP_UPDATE, 0.0, self._c_new, self._ews,
                               c_old=self.rnm_log_consts,
                               old_colsum=lazy.old_colsum if lazy is not None else self._colsum,
                               arena=arena, b=b, order=self._col_order)
        if graph is not None:
            a.knn_graph = graph.data_ptr()
            a.knn_graph_dtype = _lib.idx_code(graph.dtype)
            a.knn_k = graph.size(1)
        a.nu0 = self._nu0
        a.block, a.num_blocks, a.cap = b, arena.num_blocks, arena.cap
        a.nnz_all = arena.nnz_dev.data_ptr()
        a.cols_all = arena.col.data_ptr()
        a.vals_all = arena.val.data_ptr()
        if lazy is not None:
            a.G, a.U, a.U_old = lazy.G.data_ptr(), lazy.U.data_ptr(), lazy.U_old.data_ptr()
            a.lazy_ws = lazy.ws.data_ptr()
            a.rebal_ws = lazy.exact_ws.data_ptr()
        else:
            a.rebal_ws = self._rebal_ws.data_ptr()
        a.host_flags = self._host_flags.data_ptr()
        a.ev_flags = self._flag_event.cuda_event
        a.flag_words = self._flag_words.ptr
        a.abort_flag = arena.abort_ptr.data_ptr()
        self._step_cache["built"][b] = True
        return a

    def _run_steps(self, X, first, count, sparsity_threshold, count_steps=False):
        """Partial steps of blocks first .. first+count-1: posterior,
        partial_expectation_step and rebalance_responsibilities_
        (lakde/local_adaptive_kde.py:66-181) through the native step driver,
        with one host synchronisation per step.  ``count_steps``: add the
        completed steps to iter_steps with the rest of the bookkeeping."""
        self._count_steps = count_steps
        self._ensure_prior_cache(X)
        lazy = self._lazy
        t = float(sparsity_threshold or 0.0)
        bf16 = self.precision == "bf16"
        m32 = self._uses_m32()
        frag = self._frag_blk.data_ptr() if (bf16 or m32) else None
        sc = self._screen if (bf16 or m32) else None
        for b in range(first, first + count):
            a = self._step_args(X, b)
            a.estep.threshold = t
            a.post.prior = self._prior_ptr(b, a.post.prior)
            a.bf16_frag = frag
            a.screen_xb = sc.xb.data_ptr() if bf16 else None
            a.fp32_mfma = 1 if m32 else 0
            a.screen_xnorm = sc.xnorm.data_ptr() if sc is not None else None
            a.screen_origin = sc.origin.data_ptr() if sc is not None else None
            split = bf16 and self.screen_split
            a.screen_split_ws = self._split_ws.data_ptr() if split else None
            a.screen_split_bytes = self._split_bytes if split else 0
            ev = self._estep_timer.pair() if self._estep_timer is not None else None
            if ev is not None:
                a.ev_main[0], a.ev_main[1] = ev[0].cuda_event, ev[1].cuda_event
            else:
                a.ev_main[0] = a.ev_main[1] = None
        lib = _lib.load()
        st = stream_of(X)
        arena = self.arena
        b, end = first, first + count
        while b < end:
            # the native steps and the host bookkeeping that follows them
            # (entry counts, normaliser buffer roles, lazy state) complete
            # together: a SIGINT in between is delivered afterwards
            # (lakde/local_knn_kde.py:224 wraps its state writes the same way)
            with InterruptDelay():
                cbuf = (ctypes.c_void_p * 2)(self.rnm_log_consts.data_ptr(),
                                             self._c_new.data_ptr())
                cur = ctypes.c_int(0)
                dirty = ctypes.c_int(int(lazy is not None and lazy.dirty))
                done = ctypes.c_int64(0)
                rc = lib.lakde_partial_sweep(self._step_cache["arr"], b, end - b, cbuf,
                                             ctypes.byref(cur), ctypes.byref(dirty),
                                             ctypes.byref(done), st)
                self._account(b, b + done.value, cur.value, dirty.value)
                b += done.value
                if rc == LAKDE_OK:
                    break
                if rc == LAKDE_ESTABILIZE:
                    raise RuntimeError("Stabilization did not converge")
                if rc != LAKDE_ECAPACITY:
                    check(rc, "partial_sweep")
                # the block's entries exceed the arena: grow, rebuild the
                # arguments (same per-step fields) and finish the step
                old = self._step_cache["arr"][b]
                keep = (old.estep.threshold, old.post.prior, old.bf16_frag, old.ev_main[0],
                        old.ev_main[1], old.fold_first)
                keep_sc = (old.screen_xb, old.screen_xnorm, old.screen_origin, old.fp32_mfma,
                           old.screen_split_ws, old.screen_split_bytes)
                arena.grow(int(self._flags[b]))
                self._step_cache = {}
                for j in range(b, end):
                    a = self._step_args(X, j)
                    a.estep.threshold, a.post.prior, a.bf16_frag = keep[0], keep[1], keep[2]
                    (a.screen_xb, a.screen_xnorm, a.screen_origin, a.fp32_mfma,
                     a.screen_split_ws, a.screen_split_bytes) = keep_sc
                    if j == b:
                        a.ev_main[0], a.ev_main[1], a.fold_first = keep[3], keep[4], keep[5]
                a = self._step_cache["arr"][b]
                a.estep.c_old = self.rnm_log_consts.data_ptr()
                a.estep.c = self._c_new.data_ptr()
                check(lib.lakde_partial_step_resume(ctypes.byref(a), st), "partial_step_resume")
                unsafe = (int(self._flags[arena.num_blocks + 1]) & 1) != 0
                self._account(b, b + 1, 1, int(lazy is not None and not unsafe))
                b += 1

    def _account(self, first, stop, cur, dirty):
        """Host bookkeeping after driver steps first .. stop-1: entry counts
        (from the copied arena counts), the normaliser buffers' roles, the
        lazy state."""
        arena = self.arena
        for j in range(first, stop):
            nnz = int(self._flags[j])
            self._active += nnz - arena.nnz[j]
            arena.nnz[j] = nnz
        if cur == 1:
            self.rnm_log_consts, self._c_new = self._c_new, self.rnm_log_consts
        if self._lazy is not None and stop > first:
            self._lazy.dirty = bool(dirty)
        if self._count_steps:
            self.iter_steps += stop - first

    def _partial_step(self, X, b, sparsity_threshold):
        self._run_steps(X, b, 1, sparsity_threshold)

    def _prior_ptr(self, b, current):
        """Device address of the prior for block b's step (the shared prior
        of the hierarchical model is a new tensor every step)."""
        return current

    def _sweep_fusable(self):
        """Whether a whole sweep can run in the native loop (no host work
        between partial steps)."""
        return False

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
        """lakde/kdes.py:34-87.  Without callbacks, and where no host work
        separates the partial steps (LocalKNearestKDE), each VI iteration
        runs in the native sweep loop, in calls of at most ``sweep_chunk``
        partial steps: a KeyboardInterrupt (SIGINT) takes effect after the
        call in progress, with the bookkeeping of its steps done, and
        ``fit`` returns a consistent model, as the reference's does."""
        if callbacks is not None or not self._sweep_fusable():
            super().fit(X, iterations, callbacks, sparsity_threshold, dataset_label)
            self.materialize()
            return self
        assert X.is_contiguous(), "Training data must be contiguous, call .contiguous() before fitting"
        if sparsity_threshold is None:
            sparsity_threshold = 0
        if self.iter_steps < 1:
            self.init_parameters_(X, sparsity_threshold)
            self.init_summaries_(X, sparsity_threshold)
        num_blocks = (X.size(0) - 1) // self.block_size + 1
        chunk = max(1, int(self.sweep_chunk))
        try:
            for _ in range(iterations):
                for b0 in range(0, num_blocks, chunk):
                    self._run_steps(X, b0, min(chunk, num_blocks - b0), sparsity_threshold,
                                    count_steps=True)
        except (KeyboardInterrupt, StopIteration):
            pass
        finally:
            self._count_steps = False
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
            if self.pred_screen:
                if self._screen is None:
                    self._screen = F.ScreenPoints(X)
                frag = torch.empty(N * F.screen_frag_bytes(D), dtype=torch.uint8,
                                   device=Y.device)
                F.pack_screen(self.rec, D, self._screen.origin, frag)
                pts = F.ScreenPoints(Y, origin=self._screen.origin)
                return F.pred_density_screen(self.rec, frag, Y, pts, out)
            frag = torch.empty(N * F.bf16_frag_bytes(D), dtype=torch.uint8, device=Y.device)
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
