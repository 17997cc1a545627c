"""CPU oracle: a numpy restatement of lakde's pure-PyTorch (LAKDE_NO_EXT=1) path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``lakde_amd``) may import
this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Every function is parameterised by a numpy dtype ``dt``: ``np.float32``
mirrors the reference's FP32 arithmetic (including FP32 jitter decisions with
``finfo(float32).eps``), ``np.float64`` is the FP64 oracle.  Each function cites
the reference file:line it restates.  Deliberate differences from the
reference, all order-only:

* responsibilities are kept in canonical CSR order (row ascending, column
  ascending within a row); the reference sorts rows with a non-stable argsort
  (lakde/local_adaptive_kde.py:149, lakde/sparse_util.py:17), which fixes only
  the row order;
* scatter-adds run through ``np.bincount`` (float64 accumulation, then cast).

How each part is pinned (checked independently of this file) is listed in the
tests: closed forms and brute force for the quad form, accumulation, ind2ptr,
Cholesky/jitter and the Student-t predictive density (scipy), and golden
vectors written by running the reference itself on small inputs for the full
fit trajectory (tests/golden/).
"""

import math

import numpy as np
from scipy import special

__all__ = [
    "ind2ptr", "chol_quad_form", "calc_W_inv", "lgamma", "mvlgamma", "mvdigamma",
    "chol_logdet", "chol_mm_trace", "logsubexp", "logaddexp", "logsumexp0",
    "tril_inverse", "cholesky_ex", "stabilize", "stabilized_cholesky",
    "batched_knn_cov", "compute_knn_graph", "graph_to_slices", "cov",
    "calc_log_rho_block", "LocalKNearestKDE", "LocalHierarchicalKDE",
]


# ----------------------------------------------------------------- ops (L1)

def ind2ptr(row, M):
    """rowptr[i] = #{e : row[e] < i}, i = 0..M (lakde/kernels.py:11-17).

    The reference counts runs with unique_consecutive and cumsums; for sorted
    ``row`` that equals a lower-bound search, which is what this computes.
    """
    row = np.asarray(row)
    return np.searchsorted(row, np.arange(M + 1), side="left").astype(row.dtype)


def chol_quad_form(W_triu, Xn, Xm, dt=None, chunk=4096):
    """Q[n, m] = ||W_triu[n]^T (Xm[m] - Xn[n])||^2 (lakde/kernels.py:20-23)."""
    dt = dt or W_triu.dtype
    W_triu = np.asarray(W_triu, dt)
    Xn = np.asarray(Xn, dt)
    Xm = np.asarray(Xm, dt)
    out = np.empty((Xn.shape[0], Xm.shape[0]), dt)
    Lt = np.swapaxes(W_triu, 1, 2)  # lower triangular U^T
    for s in range(0, Xm.shape[0], chunk):
        diff = Xm[None, s:s + chunk] - Xn[:, None]           # (Nb, Mc, D)
        LX = np.einsum("nij,nmj->nmi", Lt, diff, optimize=True)
        out[:, s:s + chunk] = np.sum(LX * LX, axis=-1)
    return out


def calc_W_inv(rowptr, col, value, Xn, Xm, dt=None):
    """S_n = sum_{e in row n} value[e] (Xm[col[e]] - Xn[n]) (...)^T
    (lakde/kernels.py:26-39)."""
    dt = dt or value.dtype
    Xn = np.asarray(Xn, dt)
    Xm = np.asarray(Xm, dt)
    value = np.asarray(value, dt)
    N, D = Xn.shape
    rowptr = np.asarray(rowptr, np.int64)
    out = np.zeros((N, D, D), dt)
    lo, hi = int(rowptr[0]), int(rowptr[-1])
    if hi <= lo:
        return out
    counts = np.diff(rowptr)
    rows = np.repeat(np.arange(N), counts)
    c = np.asarray(col[lo:hi], np.int64)
    diff = Xm[c] - Xn[rows]
    outer = (value[lo:hi, None, None] * diff[:, :, None]) * diff[:, None, :]
    nz = counts > 0
    starts = rowptr[:-1][nz] - lo
    out[nz] = np.add.reduceat(outer, starts, axis=0)
    return out


# ------------------------------------------------------- numerics (L2)

def lgamma(a):
    """lgamma evaluated in float64, cast back (lakde/linalg_util.py:8-9)."""
    a = np.asarray(a)
    return special.gammaln(a.astype(np.float64)).astype(a.dtype)


def mvlgamma(a, D):
    """lakde/linalg_util.py:13-18."""
    a = np.asarray(a)
    i = np.arange(1, D + 1, dtype=a.dtype)
    coeff = D * (D - 1) / 4 * math.log(math.pi)
    return coeff + np.sum(lgamma(a[..., None] + (1 - i) / 2), axis=-1)


def mvdigamma(a, D):
    """sum_i digamma(a + (1 - i)/2), i = 1..D (lakde/linalg_util.py:22-26)."""
    a = np.asarray(a)
    i = np.arange(1, D + 1, dtype=a.dtype)
    return np.sum(special.digamma(a[..., None] + (1 - i) / 2), axis=-1).astype(a.dtype)


def chol_logdet(L):
    """2 sum log diag(L) (lakde/linalg_util.py:30-31)."""
    return 2 * np.log(np.diagonal(L, axis1=-2, axis2=-1)).sum(axis=-1)


def chol_mm_trace(L, R):
    """sum((L^T R)^2) (lakde/linalg_util.py:35-37)."""
    LR = np.swapaxes(L, -1, -2) @ R
    return np.sum(LR * LR, axis=(-1, -2))


def logsubexp(a, b):
    """a + log1p(max(-exp(b - a), -1)) (lakde/linalg_util.py:41-42)."""
    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
        return a + np.log1p(np.maximum(-np.exp(b - a), -1))


def logaddexp(a, b):
    """torch.logaddexp semantics (both -inf -> -inf)."""
    with np.errstate(invalid="ignore"):
        return np.logaddexp(a, b)


def logsumexp0(x):
    """torch.logsumexp(x, dim=0) with an all -inf column giving -inf."""
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        mx = np.max(x, axis=0)
        safe = np.where(np.isfinite(mx), mx, 0).astype(x.dtype)
        s = np.sum(np.exp(x - safe[None]), axis=0)
        return (np.log(s) + safe).astype(x.dtype)


def tril_inverse(L):
    """Forward substitution L X = I (lakde/linalg_util.py:49-52)."""
    L = np.asarray(L)
    B = L[None] if L.ndim == 2 else L
    D = B.shape[-1]
    X = np.zeros_like(B)
    eye = np.eye(D, dtype=B.dtype)
    for i in range(D):
        acc = eye[i][None, :] - np.einsum("bk,bkj->bj", B[:, i, :i], X[:, :i, :])
        X[:, i, :] = acc / B[:, i, i, None]
    return X[0] if L.ndim == 2 else X


def cholesky_ex(A):
    """Batched lower Cholesky with per-matrix failure flags (LAPACK potrf
    semantics, as torch.linalg.cholesky_ex).  Returns (L, fail[bool])."""
    A = np.asarray(A)
    single = A.ndim == 2
    B = A[None] if single else A
    L = np.zeros_like(B)
    fail = np.zeros(B.shape[0], bool)
    try:
        L[:] = np.linalg.cholesky(B)
    except np.linalg.LinAlgError:
        for i in range(B.shape[0]):
            try:
                L[i] = np.linalg.cholesky(B[i])
            except np.linalg.LinAlgError:
                fail[i] = True
    bad = ~np.all(np.isfinite(L.reshape(L.shape[0], -1)), axis=1)
    fail |= bad
    if single:
        return L[0], fail[0]
    return L, fail


def _add_diag(X, j):
    out = X.copy()
    D = X.shape[-1]
    idx = np.arange(D)
    out[..., idx, idx] = out[..., idx, idx] + np.asarray(j, X.dtype)[..., None]
    return out


def stabilize(X):
    """Cumulative jitter eps*2^k, k = 0..31, until every matrix factors
    (lakde/linalg_util.py:109-126).  Returns the per-matrix jitter."""
    eps = np.finfo(X.dtype).eps
    jitter = np.zeros(X.shape[:-2], X.dtype)
    mask = np.ones(X.shape[:-2], bool)
    for k in range(32):
        jitter[mask] = (jitter[mask] + X.dtype.type(eps * 2 ** k)).astype(X.dtype)
        _, mask = cholesky_ex(_add_diag(X, jitter))
        if not np.any(mask):
            return jitter
    raise RuntimeError("Stabilization did not converge")


def stabilized_cholesky(X, jitter=None):
    """Persistent-jitter Cholesky (lakde/linalg_util.py:129-170).

    If any stored jitter is nonzero the factorisation is of X + diag(jitter);
    matrices that still fail get a fresh jitter from ``stabilize`` on X (not on
    X + old jitter), then everything is refactored.  Returns (L, jitter).
    """
    X = np.asarray(X)
    if jitter is not None and np.count_nonzero(jitter) != 0:
        L, fail = cholesky_ex(_add_diag(X, jitter))
    else:
        L, fail = cholesky_ex(X)
    if not np.any(fail):
        return L, jitter
    if jitter is None:
        jitter = np.zeros(X.shape[:-2], X.dtype)
    else:
        jitter = np.array(jitter, copy=True)
    if X.ndim == 2:
        jitter = np.asarray(stabilize(X[None])[0], jitter.dtype)
    else:
        jitter[fail] = stabilize(X[fail])
    return stabilized_cholesky(X, jitter)


def _shared_cholesky(X, jit32):
    """stabilized_cholesky on the FP64 shared precision with the reference's
    FP32 0-dim jitter (lakde/local_hierarchical_kde.py:307-309 calling
    lakde/linalg_util.py:129-170): the jitter found with FP64 eps is stored
    rounded to the jitter's own dtype before the refactorisation."""
    j = float(jit32)
    L, fail = cholesky_ex(_add_diag(X, j) if j != 0 else X)
    if not fail:
        return L, jit32
    j32 = type(jit32)(stabilize(X[None])[0])
    L, fail = cholesky_ex(_add_diag(X, float(j32)))
    if fail:
        raise RuntimeError("shared prior refactorisation failed after stabilisation")
    return L, j32


def cov(X):
    """Sample covariance in float64 (lakde/utils.py:28-37)."""
    X = np.asarray(X, np.float64)
    Xc = X - X.mean(axis=0, keepdims=True)
    return (Xc.T @ Xc) / (X.shape[0] - 1)


def batched_knn_cov(X, X_self, knn_graph, dt):
    """Self-centred kNN covariance, divided by K (lakde/utils.py:40-65 with
    include_self=False, self_center=True)."""
    X = np.asarray(X, dt)
    X_self = np.asarray(X_self, dt)
    M, K = knn_graph.shape
    Xc = X[np.asarray(knn_graph, np.int64)] - X_self[:, None]   # (M, K, D)
    out = np.einsum("mkd,mke->mde", Xc, Xc, optimize=True)
    return (out / dt(K)).astype(dt)


def compute_knn_graph(X, k, include_self=False, bsize=1024):
    """k nearest neighbours by exact L2 distance, ties broken by index
    (lakde/utils.py:77-116; the reference's torch.topk tie order is
    implementation-defined)."""
    X64 = np.asarray(X, np.float64)
    N = X64.shape[0]
    out = np.empty((N, k), np.int32)
    sq = np.sum(X64 * X64, axis=1)
    for i in range(0, N, bsize):
        xt = X64[i:i + bsize]
        d2 = sq[i:i + bsize, None] + sq[None] - 2 * xt @ X64.T
        d2 = np.maximum(d2, 0)
        # refine with exact differences on the candidates to avoid mm rounding
        if not include_self:
            r = np.arange(xt.shape[0])
            d2[r, i + r] = np.inf
        cand = np.argpartition(d2, min(k + 8, N - 1), axis=1)[:, :min(k + 9, N)]
        exact = np.sum((X64[cand] - xt[:, None]) ** 2, axis=-1)
        if not include_self:
            exact[cand == (i + np.arange(xt.shape[0]))[:, None]] = np.inf
        order = np.lexsort((cand, exact), axis=1)[:, :k]
        out[i:i + bsize] = np.take_along_axis(cand, order, axis=1)
    return out


def graph_to_slices(graph, block_size):
    """kNN graph -> per kernel-row-block COO (row local, col global), in
    canonical (row, col) order (lakde/sparse_util.py:8-54).

    Entry (row=graph[q, j], col=q) for every query q and neighbour slot j.
    """
    N, K = graph.shape
    row = np.asarray(graph, np.int64).reshape(-1)
    col = np.repeat(np.arange(N, dtype=np.int64), K)
    order = np.lexsort((col, row))
    row, col = row[order], col[order]
    rowptr = ind2ptr(row, N)
    slices = []
    for i in range(0, N, block_size):
        a, b = rowptr[i], rowptr[min(i + block_size, N)]
        slices.append(np.stack([row[a:b] - i, col[a:b]]).astype(np.int32))
    return slices


def calc_log_rho_block(X_test, X_train, nu, W_triu, logdet_W, dt):
    """log rho = -nu/2 Q + (mvdigamma(nu/2, D) + D ln 2 + ln|W|)/2
    (lakde/local_adaptive_kde.py:36-44)."""
    D = X_test.shape[-1]
    expect_logdet_lambda = (mvdigamma(nu / 2, D) + dt(D * math.log(2)) + logdet_W).astype(dt)
    exponent = (dt(-0.5) * nu).astype(dt)
    const = (dt(0.5) * expect_logdet_lambda).astype(dt)
    Q = chol_quad_form(W_triu, X_train, X_test, dt=dt)
    Q *= exponent[:, None]
    Q += const[:, None]
    return Q


# ------------------------------------------------------------ VI core (L3)

def _as_np(x):
    """A torch tensor (any device) or array-like as a numpy array."""
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.asarray(x)


class _LocallyAdaptiveKDE:
    """lakde/local_adaptive_kde.py:47-223 (plus the fit driver of
    lakde/kdes.py:34-87)."""

    def __init__(self, nu_0, a_0, b_0, block_size, dtype):
        self.nu_0 = nu_0
        self.a_0 = a_0
        self.b_0 = b_0
        self.block_size = block_size
        self.dt = np.dtype(dtype).type
        self.iter_steps = 0

    # lakde/local_adaptive_kde.py:66-81
    def calc_posterior_contribs(self, X, block_idx):
        s = block_idx * self.block_size
        X_train = X[s:s + self.block_size]
        indices, values = self.rnm_blocks[block_idx]
        row, col = indices
        rowptr = ind2ptr(row.astype(np.int64), X_train.shape[0])
        W_inv_contrib = calc_W_inv(rowptr, col, values, X_train, X, dt=self.dt)
        nu = (self.nu_0 + np.bincount(row, weights=values,
                                      minlength=X_train.shape[0])).astype(self.dt)
        return W_inv_contrib, nu

    # lakde/local_adaptive_kde.py:83-96
    def calc_posterior(self, block_idx, sigma_0, W_inv_contrib):
        dt = self.dt
        s = block_idx * self.block_size
        e = s + self.block_size
        expect_tau = (self.a[s:e] / self.b[s:e]).astype(dt)
        scale = (dt(self.nu_0) * expect_tau).astype(dt)
        W_inv = (np.asarray(sigma_0).astype(dt) * scale[:, None, None]).astype(dt)
        W_inv += W_inv_contrib
        L, jit = stabilized_cholesky(W_inv, self.W_inv_jitter[s:e])
        self.W_inv_jitter[s:e] = jit
        return np.swapaxes(tril_inverse(L), -1, -2)

    # lakde/local_adaptive_kde.py:98-154
    def partial_expectation_step(self, X, block_idx, W_triu, nu,
                                 update_normalization=True, sparsity_threshold=None):
        dt = self.dt
        N, D = X.shape
        bs = self.block_size
        s = block_idx * bs
        X_train = X[s:s + bs]
        logdet_W = chol_logdet(W_triu).astype(dt)
        old = self.rnm_log_consts
        c = old
        if update_normalization:
            c = old.copy()
            indices, values = self.rnm_blocks[block_idx]
            old_sums = np.bincount(indices[1], weights=values, minlength=N).astype(dt)
            with np.errstate(divide="ignore"):
                old_contrib = (np.log(old_sums) + old).astype(dt)
        rows, cols, vals = [], [], []
        for bj, j in enumerate(range(0, N, bs)):
            X_test = X[j:j + bs]
            log_rho = calc_log_rho_block(X_test, X_train, nu, W_triu, logdet_W, dt)
            if block_idx == bj:
                np.fill_diagonal(log_rho, -np.inf)
            if update_normalization:
                new = logsumexp0(log_rho)
                c[j:j + bs] = logaddexp(logsubexp(old[j:j + bs], old_contrib[j:j + bs]), new)
            with np.errstate(invalid="ignore"):
                rnm = np.exp(log_rho - c[j:j + bs][None])
            if sparsity_threshold:
                keep = rnm > sparsity_threshold
            else:
                keep = rnm != 0
            r, cc = np.nonzero(keep)
            rows.append(r)
            cols.append(cc + j)
            vals.append(rnm[r, cc])
        rows = np.concatenate(rows)
        cols = np.concatenate(cols)
        vals = np.concatenate(vals).astype(dt)
        order = np.lexsort((cols, rows))
        indices = np.stack([rows[order], cols[order]]).astype(np.int32)
        if update_normalization:
            return indices, vals[order], c
        return indices, vals[order]

    # lakde/local_adaptive_kde.py:156-181
    def rebalance_responsibilities_(self, X, block_idx, rnm_log_consts):
        dt = self.dt
        N = X.shape[0]
        fi = np.finfo(dt)
        with np.errstate(over="ignore", invalid="ignore"):
            coeff = np.exp(self.rnm_log_consts - rnm_log_consts).astype(dt)
        coeff = np.clip(coeff, fi.tiny, fi.max)
        sums = np.zeros(N, np.float64)
        for bi in range(len(self.rnm_blocks)):
            indices, values = self.rnm_blocks[bi]
            col = indices[1]
            if bi != block_idx:
                values *= coeff[col]
            sums += np.bincount(col, weights=values, minlength=N)
        sums = np.maximum(sums.astype(dt), fi.eps)
        rnm_log_consts = (rnm_log_consts + np.log(sums)).astype(dt)
        coeff = np.clip((dt(1) / sums).astype(dt), fi.tiny, fi.max)
        for bi in range(len(self.rnm_blocks)):
            indices, values = self.rnm_blocks[bi]
            values *= coeff[indices[1]]
        return rnm_log_consts

    # lakde/local_adaptive_kde.py:200-223
    def log_pred_density(self, X, Y):
        dt = self.dt
        X = np.asarray(X, dt)
        Y = np.asarray(Y, dt)
        N, D = X.shape
        bs = self.block_size
        acc = np.full(Y.shape[0], -np.inf, dt)
        for bi, i in enumerate(range(0, N, bs)):
            X_train = X[i:i + bs]
            contrib, nu = self.calc_posterior_contribs(X, bi)
            sigma_0, _ = self.calc_prior_cov(X, bi)
            W_triu = self.calc_posterior(bi, sigma_0, contrib)
            log_coeff = (lgamma((nu + 1) / 2) - lgamma((nu + 1 - D) / 2)
                         + dt(0.5) * chol_logdet(W_triu)
                         - dt(D / 2 * math.log(math.pi))).astype(dt)
            for j in range(0, Y.shape[0], bs):
                Q = chol_quad_form(W_triu, X_train, Y[j:j + bs], dt=dt)
                lp = log_coeff[:, None] - ((nu[:, None] + 1) / 2) * np.log1p(Q)
                acc[j:j + bs] = logaddexp(acc[j:j + bs], logsumexp0(lp.astype(dt)))
        return (dt(-math.log(N)) + acc).astype(dt)

    # lakde/kdes.py:115-125, lakde/local_adaptive_kde.py:272-309: the state
    # under the reference's keys (numpy arrays; rnm_blocks a list of
    # (indices (2, nnz) int32, values)).  load_state_dict takes the same
    # mapping with torch tensors or arrays, as torch.load returns the
    # reference's saved state.
    _STATE_POINT = ("nu", "a", "b", "rnm_log_consts", "logdet_W", "W_inv_jitter")

    def state_dict(self):
        d = {"block_size": self.block_size, "verbose": False, "iter_steps": self.iter_steps,
             "nu_0": self.nu_0, "a_0": self.a_0, "b_0": self.b_0,
             "rnm_blocks": [(ix.copy(), v.copy()) for ix, v in self.rnm_blocks],
             "rnm_active_contribs": self.dt(sum(v.size for _, v in self.rnm_blocks))}
        for k in self._STATE_POINT:
            d[k] = getattr(self, k).copy()
        return d

    def load_state_dict(self, d):
        dt = self.dt
        self.block_size = int(d["block_size"])
        self.iter_steps = int(d["iter_steps"])
        for k in ("nu_0", "a_0", "b_0"):
            setattr(self, k, dt(_as_np(d[k])))
        for k in self._STATE_POINT:
            setattr(self, k, np.array(_as_np(d[k]), dt))
        self.rnm_blocks = [(np.array(_as_np(ix), np.int32), np.array(_as_np(v), dt))
                           for ix, v in d["rnm_blocks"]]

    # lakde/kdes.py:34-87
    def fit(self, X, iterations=200, sparsity_threshold=None, callbacks=()):
        X = np.ascontiguousarray(X, self.dt)
        N = X.shape[0]
        if sparsity_threshold is None:
            sparsity_threshold = 0
        if self.iter_steps < 1:
            self.init_parameters_(X, sparsity_threshold)
            self.init_summaries_(X, sparsity_threshold)
        num_blocks = (N - 1) // self.block_size + 1
        for _ in range(iterations):
            for bi in range(num_blocks):
                self.partial_step_(X, bi, sparsity_threshold=sparsity_threshold)
                for cb in callbacks:
                    cb(X, self, self.iter_steps)
                self.iter_steps += 1
        return self

    def _init_common(self, X, sparsity_threshold):
        """Shared part of init_parameters_ (lakde/local_knn_kde.py:73-138,
        lakde/local_hierarchical_kde.py:111-196)."""
        dt = self.dt
        N, D = X.shape
        self.block_size = min(self.block_size, N)
        self.W_inv_jitter = np.zeros(N, dt)
        self.a_0 = dt(self.a_0)
        self.b_0 = dt(self.b_0)
        self.nu_0 = dt(D) if self.nu_0 is None else dt(self.nu_0)
        if self.nu_0 <= D - 1:
            raise ValueError("nu_0 must be greater than the data dimensionality minus one")

    def _init_rnm(self, X, graph, sparsity_threshold):
        dt = self.dt
        N, D = X.shape
        if sparsity_threshold:
            init_k = min(math.floor(1 / sparsity_threshold), graph.shape[1])
            init_graph = graph[:, :init_k]
        else:
            init_graph, init_k = graph, graph.shape[1]
        inds = graph_to_slices(init_graph, self.block_size)
        self.rnm_blocks = [(ix, np.full(ix.shape[1], dt(1.0 / init_k), dt)) for ix in inds]
        self.nu = np.empty(N, dt)
        self.a = np.full(N, dt(self.a_0 + self.nu_0 * dt(D) / dt(2)), dt)
        self.b = self.a.copy()
        self.rnm_log_consts = np.full(N, -np.inf, dt)
        self.logdet_W = np.empty(N, dt)

    def _lse_pass(self, X, bi, W_triu, nu, logdet_W):
        """lakde/local_knn_kde.py:157-166."""
        N = X.shape[0]
        bs = self.block_size
        X_train = X[bi * bs:(bi + 1) * bs]
        for bj, j in enumerate(range(0, N, bs)):
            log_rho = calc_log_rho_block(X[j:j + bs], X_train, nu, W_triu, logdet_W, self.dt)
            if bi == bj:
                np.fill_diagonal(log_rho, -np.inf)
            self.rnm_log_consts[j:j + bs] = logaddexp(self.rnm_log_consts[j:j + bs],
                                                      logsumexp0(log_rho))


class LocalKNearestKDE(_LocallyAdaptiveKDE):
    """lakde/local_knn_kde.py:50-254."""

    def __init__(self, nu_0=None, k_or_knn_graph=None, block_size=512,
                 a_0=1e-6, b_0=1e-6, dtype=np.float32):
        super().__init__(nu_0, a_0, b_0, block_size, dtype)
        self.k_or_knn_graph = k_or_knn_graph

    # lakde/local_knn_kde.py:73-138
    def init_parameters_(self, X, sparsity_threshold=None):
        self._init_common(X, sparsity_threshold)
        N = X.shape[0]
        self.sigma_0_jitter = np.zeros(N, self.dt)
        g = self.k_or_knn_graph
        if isinstance(g, (int, np.integer)):
            g = compute_knn_graph(X, int(g), include_self=False)
        self.knn_graph = np.asarray(g, np.int32)
        self.k = self.knn_graph.shape[1]
        self._init_rnm(X, self.knn_graph, sparsity_threshold)
        self.logdet_sigma_0 = np.empty(N, self.dt)
        self.sigma_0_lambda_mm_trace = np.empty(N, self.dt)

    # lakde/local_knn_kde.py:343-367
    def state_dict(self):
        d = super().state_dict()
        d.update({"knn_graph": self.knn_graph.copy(), "k": self.k,
                  "logdet_sigma_0": self.logdet_sigma_0.copy(),
                  "sigma_0_lambda_mm_trace": self.sigma_0_lambda_mm_trace.copy(),
                  "sigma_0_jitter": self.sigma_0_jitter.copy()})
        return d

    def load_state_dict(self, d):
        super().load_state_dict(d)
        self.knn_graph = np.array(_as_np(d["knn_graph"]), np.int32)
        self.k = int(d["k"])
        for k in ("logdet_sigma_0", "sigma_0_lambda_mm_trace", "sigma_0_jitter"):
            setattr(self, k, np.array(_as_np(d[k]), self.dt))

    # lakde/local_knn_kde.py:195-207
    def calc_prior_cov(self, X, block_idx):
        s = block_idx * self.block_size
        e = s + self.block_size
        sigma_0 = batched_knn_cov(X, X[s:e], self.knn_graph[s:e], self.dt)
        L, jit = stabilized_cholesky(sigma_0, self.sigma_0_jitter[s:e])
        self.sigma_0_jitter[s:e] = jit
        sigma_0 = _add_diag(sigma_0, self.sigma_0_jitter[s:e])
        return sigma_0, L

    # lakde/local_knn_kde.py:140-193
    def init_summaries_(self, X, sparsity_threshold=None):
        dt = self.dt
        N = X.shape[0]
        bs = self.block_size
        for bi, i in enumerate(range(0, N, bs)):
            sigma_0, sigma_0_tril = self.calc_prior_cov(X, bi)
            contrib, nu = self.calc_posterior_contribs(X, bi)
            W_triu = self.calc_posterior(bi, sigma_0, contrib)
            logdet_W = chol_logdet(W_triu).astype(dt)
            self.nu[i:i + bs] = nu
            self.logdet_W[i:i + bs] = logdet_W
            self.logdet_sigma_0[i:i + bs] = chol_logdet(sigma_0_tril)
            self.sigma_0_lambda_mm_trace[i:i + bs] = nu * chol_mm_trace(sigma_0_tril, W_triu)
            self._lse_pass(X, bi, W_triu, nu, logdet_W)
        for bi, i in enumerate(range(0, N, bs)):
            sigma_0, _ = self.calc_prior_cov(X, bi)
            contrib, nu = self.calc_posterior_contribs(X, bi)
            W_triu = self.calc_posterior(bi, sigma_0, contrib)
            self.rnm_blocks[bi] = self.partial_expectation_step(
                X, bi, W_triu, nu, update_normalization=False,
                sparsity_threshold=sparsity_threshold)
        self.rnm_log_consts = self.rebalance_responsibilities_(X, -1, self.rnm_log_consts)

    # lakde/local_knn_kde.py:209-239
    def partial_step_(self, X, block_idx, sparsity_threshold=None):
        dt = self.dt
        s = block_idx * self.block_size
        e = s + self.block_size
        sigma_0, sigma_0_tril = self.calc_prior_cov(X, block_idx)
        contrib, nu = self.calc_posterior_contribs(X, block_idx)
        W_triu = self.calc_posterior(block_idx, sigma_0, contrib)
        ind, val, c = self.partial_expectation_step(
            X, block_idx, W_triu, nu, sparsity_threshold=sparsity_threshold)
        logdet_W = chol_logdet(W_triu).astype(dt)
        logdet_sigma_0 = chol_logdet(sigma_0_tril).astype(dt)
        tr = (nu * chol_mm_trace(sigma_0_tril, W_triu)).astype(dt)
        self.rnm_blocks[block_idx] = (ind, val)
        self.logdet_W[s:e] = logdet_W
        self.logdet_sigma_0[s:e] = logdet_sigma_0
        self.sigma_0_lambda_mm_trace[s:e] = tr
        self.nu[s:e] = nu
        self.b[s:e] = self.b_0 + self.nu_0 * tr / dt(2)
        self.rnm_log_consts = self.rebalance_responsibilities_(X, block_idx, c)


class LocalHierarchicalKDE(_LocallyAdaptiveKDE):
    """lakde/local_hierarchical_kde.py:77-332."""

    def __init__(self, nu_0=None, nu_shared_0=None, k_or_knn_graph=None,
                 block_size=512, a_0=1e-6, b_0=1e-6, dtype=np.float32):
        super().__init__(nu_0, a_0, b_0, block_size, dtype)
        self.nu_shared_0 = nu_shared_0
        self.k_or_knn_graph = k_or_knn_graph

    # lakde/local_hierarchical_kde.py:111-196
    def init_parameters_(self, X, sparsity_threshold=None):
        dt = self.dt
        self._init_common(X, sparsity_threshold)
        N, D = X.shape
        self.W_shared_inv_jitter = dt(0)
        self.nu_shared_0 = dt(D - 1 + 1e-2) if self.nu_shared_0 is None else dt(self.nu_shared_0)
        if self.nu_shared_0 <= D - 1:
            raise ValueError("nu_shared_0 must be greater than the data dimensionality minus one")
        sigma_0 = cov(X)
        sigma_0_tril, _ = stabilized_cholesky(sigma_0)
        self.lambda_0_triu = tril_inverse(sigma_0_tril).T
        self.lambda_0 = self.lambda_0_triu @ self.lambda_0_triu.T
        g = self.k_or_knn_graph
        if isinstance(g, (int, np.integer)):
            g = compute_knn_graph(X, int(g), include_self=True)
        graph = np.asarray(g, np.int32)
        self.k = graph.shape[1]
        self._init_rnm(X, graph, sparsity_threshold)
        self.nu_shared = np.float64(self.nu_shared_0) + N * np.float64(self.nu_0)
        self.W_shared_triu = np.float64(self.nu_0) * sigma_0_tril.T / self.nu_shared
        num_blocks = (N - 1) // self.block_size + 1
        self.expect_lambda_contribs = np.empty((num_blocks, D, D), np.float64)

    # lakde/local_hierarchical_kde.py:463-490 (the shared-prior state keeps
    # the dtypes it has: float64 in the reference)
    _STATE_SHARED = ("nu_shared", "lambda_0", "lambda_0_triu", "W_shared_triu",
                     "expect_lambda_contribs")

    def state_dict(self):
        d = super().state_dict()
        d.update({"k": self.k, "nu_shared_0": self.nu_shared_0,
                  "W_shared_inv_jitter": self.W_shared_inv_jitter})
        for k in self._STATE_SHARED:
            d[k] = np.array(getattr(self, k))
        return d

    def load_state_dict(self, d):
        super().load_state_dict(d)
        self.k = int(d["k"])
        self.nu_shared_0 = self.dt(_as_np(d["nu_shared_0"]))
        self.W_shared_inv_jitter = _as_np(d["W_shared_inv_jitter"])[()]
        for k in self._STATE_SHARED:
            setattr(self, k, np.array(_as_np(d[k]), np.float64))
        self.nu_shared = np.float64(self.nu_shared)

    # lakde/local_hierarchical_kde.py:269-271
    def calc_prior_cov(self, *_):
        t = math.sqrt(self.nu_shared) * self.W_shared_triu
        return t @ t.T, t

    # lakde/local_hierarchical_kde.py:260-267
    def calc_prior_prec(self):
        return (self.expect_lambda_contribs.sum(axis=0) * np.float64(self.nu_0)
                + np.float64(self.nu_shared_0) * self.lambda_0)

    def _lambda_contrib(self, W_triu, nu, expect_tau):
        dt = self.dt
        c = (W_triu @ np.swapaxes(W_triu, -1, -2)).astype(dt)
        c *= (nu * expect_tau).astype(dt)[:, None, None]
        return c.sum(axis=0, dtype=np.float64)

    # lakde/local_hierarchical_kde.py:198-258
    def init_summaries_(self, X, sparsity_threshold=None):
        dt = self.dt
        N = X.shape[0]
        bs = self.block_size
        sigma_bar, _ = self.calc_prior_cov()
        for bi, i in enumerate(range(0, N, bs)):
            contrib, nu = self.calc_posterior_contribs(X, bi)
            W_triu = self.calc_posterior(bi, sigma_bar, contrib)
            logdet_W = chol_logdet(W_triu).astype(dt)
            self.nu[i:i + bs] = nu
            self.logdet_W[i:i + bs] = logdet_W
            expect_tau = (self.a[i:i + bs] / self.b[i:i + bs]).astype(dt)
            self.expect_lambda_contribs[bi] = self._lambda_contrib(W_triu, nu, expect_tau)
            self._lse_pass(X, bi, W_triu, nu, logdet_W)
        for bi, i in enumerate(range(0, N, bs)):
            contrib, nu = self.calc_posterior_contribs(X, bi)
            W_triu = self.calc_posterior(bi, sigma_bar, contrib)
            self.rnm_blocks[bi] = self.partial_expectation_step(
                X, bi, W_triu, nu, update_normalization=False,
                sparsity_threshold=sparsity_threshold)
        self.rnm_log_consts = self.rebalance_responsibilities_(X, -1, self.rnm_log_consts)

    # lakde/local_hierarchical_kde.py:273-311
    def partial_step_(self, X, block_idx, sparsity_threshold=None):
        dt = self.dt
        s = block_idx * self.block_size
        e = s + self.block_size
        sigma_bar, _ = self.calc_prior_cov()
        contrib, nu = self.calc_posterior_contribs(X, block_idx)
        W_triu = self.calc_posterior(block_idx, sigma_bar, contrib)
        ind, val, c = self.partial_expectation_step(
            X, block_idx, W_triu, nu, sparsity_threshold=sparsity_threshold)
        logdet_W = chol_logdet(W_triu).astype(dt)
        tr = (nu * chol_mm_trace(sigma_bar.astype(dt), W_triu)).astype(dt)
        a = self.a[s:e]
        b = (self.b_0 + self.nu_0 * tr / dt(2)).astype(dt)
        expect_tau = (a / b).astype(dt)
        lam = self._lambda_contrib(W_triu, nu, expect_tau)
        self.rnm_blocks[block_idx] = (ind, val)
        self.expect_lambda_contribs[block_idx] = lam
        W_shared_inv = self.calc_prior_prec()
        L, self.W_shared_inv_jitter = _shared_cholesky(W_shared_inv, self.W_shared_inv_jitter)
        self.W_shared_triu = tril_inverse(L).T
        self.logdet_W[s:e] = logdet_W
        self.nu[s:e] = nu
        self.b[s:e] = b
        self.rnm_log_consts = self.rebalance_responsibilities_(X, block_idx, c)
