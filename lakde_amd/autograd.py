"""Differentiation form: a differentiable log_pred_density.

``mixture_log_density(Y, Linv, Xc, nu)`` is the Student-t mixture of
lakde/local_adaptive_kde.py:200-223 given explicit posterior factors
(Linv = L^{-1}, lower, i.e. the reference's W_triu transposed), as a
``torch.autograd.Function`` whose forward is the streaming predictive kernel
and whose backward is the hand-written HIP backward (gradients w.r.t. Y, Linv,
Xc and nu; the mixture weights are recomputed from the saved per-point
log-sum-exp, nothing of size N x M is kept).

``log_pred_density(model, X, Y)`` builds the factors from a fitted model so
that gradients flow to ``Y`` and, for LocalHierarchicalKDE, to the shared
prior factor: the D x D chain W_inv = prior (nu0 E[tau]) + S_n -> Cholesky
-> inverse runs in float64 torch (``hierarchical_factors``; S_n, the
responsibility scatter, comes from the HIP accumulation kernel and is held
fixed, as are the persistent jitters), and its factors feed the HIP forward
and backward.  ``prior_scale_objective`` (LocalKNearestKDE, full scale)
keeps the HIP posterior kernel's factors and differentiates the chain with
the analytic Cholesky VJP in float64 (``FactorInverse``).
"""

import torch

from . import _fused as F
from . import _lib
from ._lib import check, ptr, require_gpu, stream_of


def _pack_records(Linv, Xc, nu):
    N, D, _ = Linv.shape
    rec = torch.empty(N, F.record_floats(D), device=Linv.device)
    check(_lib.load().lakde_pack_records(ptr(Linv), ptr(Xc), ptr(nu), N, D, ptr(rec),
                                         stream_of(Linv)), "pack_records")
    return rec


def _backward(rec, Y, out, g):
    N = rec.size(0)
    M, D = Y.shape
    dev = Y.device
    lib = _lib.load()
    nbytes = int(lib.lakde_pred_backward_workspace(N, M, D))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dY = torch.empty(M, D, device=dev)
    S = torch.empty(N, D, D, device=dev)
    s1 = torch.empty(N, D, device=dev)
    beta = torch.empty(N, device=dev)
    tau = torch.empty(N, device=dev)
    check(lib.lakde_pred_backward(ptr(rec), N, ptr(Y), M, D, ptr(out), ptr(g), ptr(dY), ptr(S),
                                  ptr(s1), ptr(beta), ptr(tau), ptr(ws), nbytes, stream_of(Y)),
          "pred_backward")
    return dY, S, s1, beta, tau


class MixtureLogDensity(torch.autograd.Function):
    """Forward and backward run in float32 on the HIP kernels; inputs of
    another floating dtype (float64 factors of the differentiable
    hierarchical chain) are rounded to float32 on the way in and their
    gradients returned in their own dtype."""

    @staticmethod
    def forward(ctx, Y, Linv, Xc, nu):
        require_gpu(Y, Linv, Xc, nu)
        ctx.dtypes = (Y.dtype, Linv.dtype, Xc.dtype, nu.dtype)
        Y = Y.detach().float().contiguous()
        Linv = Linv.detach().float().contiguous()
        Xc = Xc.detach().float().contiguous()
        nu = nu.detach().float().contiguous()
        rec = _pack_records(Linv, Xc, nu)
        out = F.pred_density(rec, Y, torch.empty(Y.size(0), device=Y.device))
        ctx.save_for_backward(Y, Linv, nu, rec, out)
        return out

    @staticmethod
    def backward(ctx, g):
        Y, Linv, nu, rec, out = ctx.saved_tensors
        D = Y.size(1)
        dY, S, s1, beta, tau = _backward(rec, Y, out, g.detach().float().contiguous())
        dLinv = dXc = dnu = None
        if ctx.needs_input_grad[1]:
            diag = torch.diagonal(Linv, dim1=-2, dim2=-1)
            dLinv = torch.tril(2 * torch.bmm(Linv, S)) + torch.diag_embed(beta[:, None] / diag)
        if ctx.needs_input_grad[2]:
            dXc = -2 * torch.bmm(Linv.mT.contiguous(), torch.bmm(Linv, s1[..., None]))[..., 0]
        if ctx.needs_input_grad[3]:
            h = (nu + 1) / 2
            dnu = 0.5 * (beta * (torch.digamma(h) - torch.digamma(h - D / 2)) - tau)
        grads = [(dY if ctx.needs_input_grad[0] else None), dLinv, dXc, dnu]
        return tuple(None if gr is None else gr.to(dt) for gr, dt in zip(grads, ctx.dtypes))


def mixture_log_density(Y, Linv, Xc, nu):
    """log p(y) = -ln N + LSE_n log St(y | x_n, Linv_n, nu_n) for explicit
    factors; differentiable w.r.t. all four arguments."""
    return MixtureLogDensity.apply(Y, Linv, Xc, nu)


class FactorInverse(torch.autograd.Function):
    """Linv = chol(A)^{-1} with the forward value supplied (the HIP posterior
    kernel's factor, so the differentiable path evaluates exactly the model's
    posterior) and the analytic VJP in float64: with L = Linv^{-1},
    gL = -tril(Linv^T gLinv Linv^T), P = Phi(L^T gL) (lower triangle, half
    diagonal), gA = sym(Linv^T P Linv) (the Cholesky VJP of Murray, 2016,
    "Differentiation of the Cholesky decomposition")."""

    @staticmethod
    def forward(ctx, A, Linv_value):
        ctx.save_for_backward(Linv_value)
        return Linv_value.clone()

    @staticmethod
    def backward(ctx, gLinv):
        (Linv,) = ctx.saved_tensors
        return _factor_inverse_vjp(Linv.double(), gLinv.double())[0], None


def _factor_inverse_vjp(Li, gLi):
    """(gA, L) for Linv = chol(A)^{-1}: the float64 VJP of FactorInverse and
    the Cholesky factor L = Linv^{-1} it computes on the way."""
    D = Li.size(-1)
    eye = torch.eye(D, dtype=torch.float64, device=Li.device).expand_as(Li)
    L = torch.linalg.solve_triangular(Li, eye, upper=False)
    gL = -torch.tril(Li.mT @ gLi @ Li.mT)
    P = torch.tril(L.mT @ gL)
    P = P - 0.5 * torch.diag_embed(torch.diagonal(P, dim1=-2, dim2=-1))
    S = Li.mT @ P @ Li
    return 0.5 * (S + S.mT), L


def _scatter(model, X, b):
    """Posterior contributions S_n, nu_n of block b from the HIP kernel."""
    N, D = X.shape
    s, e, nb = model._span(b, N)
    arena = model.arena
    S = torch.empty(nb, D, D, device=X.device)
    nu = torch.empty(nb, device=X.device)
    F.posterior_contribs(arena.rowptr[b, :nb + 1], arena.col[b], arena.val[b], X[s:e], X,
                         model._nu0, S, nu, colscale=model._colscale())
    return S, nu


def log_pred_density(model, X, Y, prior_factor=None, keep=None):
    """Differentiable held-out log density of a fitted model.

    For LocalHierarchicalKDE, ``prior_factor`` (D x D, lower triangular,
    float64) is the shared prior factor W_shared_triu^T; pass a tensor with
    ``requires_grad=True`` to differentiate through it (default: the model's
    current factor, no gradient).  ``keep`` restricts the mixture to a
    subset of the kernels (see ``hierarchical_factors``; the normalisation
    is then -log of the kept count).  For LocalKNearestKDE the kNN prior is
    data and only Y (and the factors returned by ``posterior_factors``)
    carry gradients.
    """
    require_gpu(X, Y)
    from .local_hierarchical_kde import LocalHierarchicalKDE
    if isinstance(model, LocalHierarchicalKDE):
        if prior_factor is None:
            prior_factor = model.W_shared_triu.T
        Linv, nu = hierarchical_factors(model, X, prior_factor, keep)
        Xk = X if keep is None else X[keep].contiguous()
        return mixture_log_density(Y, Linv, Xk, nu)
    if keep is not None:
        raise ValueError("keep applies to LocalHierarchicalKDE only")
    Linv, nu = posterior_factors(model, X)
    return mixture_log_density(Y, Linv, X, nu)


def hierarchical_factors(model, X, prior_factor, keep=None):
    """(Linv, nu) of every kernel of a fitted LocalHierarchicalKDE as a
    differentiable function of the shared prior factor P = W_shared_triu^T
    (D x D lower, float64): the chain of lakde/local_hierarchical_kde.py:269-271
    (Sigma_bar = nu_shared W_shared_triu W_shared_triu^T) and
    lakde/local_adaptive_kde.py:83-96 (W_inv_n = Sigma_bar nu0 E[tau_n] + S_n
    + j_n I, L_n = chol(W_inv_n), Linv_n = L_n^{-1}) evaluated in float64
    torch on D x D matrices, so that its gradient w.r.t. P is that of the
    float64 chain itself.  S_n and nu_n come from the HIP accumulation kernel
    and are held fixed with E[tau_n] and the persistent jitters j_n (the
    fitted values).  Linv is float64; the predictive kernels take it rounded
    to float32.

    ``keep`` (optional bool mask of N): only those kernels' factors are
    formed and returned (with their nu), e.g. to leave out posteriors whose
    float64 chain does not factor (``unfactorable``); without it, such a
    posterior raises."""
    N, D = X.shape
    T = torch.sqrt(model.nu_shared.double()) * prior_factor.T
    sigma_bar = T @ T.T
    eye = torch.eye(D, dtype=torch.float64, device=X.device)
    Linvs, nus = [], []
    for b in model._blocks(N):
        s, e, nb = model._span(b, N)
        S, nu = _scatter(model, X, b)
        W_inv = _hier_w_inv(model, sigma_bar, S, s, e, eye)
        if keep is not None:
            k = keep[s:e]
            W_inv, nu = W_inv[k], nu[k]
        L, info = torch.linalg.cholesky_ex(W_inv)
        bad = int(torch.count_nonzero(info))
        if bad:
            raise RuntimeError(f"{bad} posterior scale matrices of block {b} do not factor "
                               "in float64")
        Linvs.append(torch.linalg.solve_triangular(L, eye.expand(W_inv.size(0), D, D),
                                                   upper=False))
        nus.append(nu)
    return torch.cat(Linvs), torch.cat(nus)


def _hier_w_inv(model, sigma_bar, S, s, e, eye):
    scale = model._nu0 * (model.a[s:e] / model.b[s:e])  # float32, as the posterior kernel
    W_inv = sigma_bar[None] * scale.double()[:, None, None] + S.double()
    return W_inv + model.W_inv_jitter[s:e].double()[:, None, None] * eye


def unfactorable(model, X, prior_factor=None):
    """Bool mask of the kernels of a fitted LocalHierarchicalKDE whose float64
    posterior chain (``hierarchical_factors``) does not factor: posteriors
    that the float32 kernel factored only with a persistent jitter of the
    size of float32 rounding can be indefinite when W_inv is formed in
    float64."""
    N, D = X.shape
    if prior_factor is None:
        prior_factor = model.W_shared_triu.T
    with torch.no_grad():
        T = torch.sqrt(model.nu_shared.double()) * prior_factor.T
        sigma_bar = T @ T.T
        eye = torch.eye(D, dtype=torch.float64, device=X.device)
        out = []
        for b in model._blocks(N):
            s, e, nb = model._span(b, N)
            S, _ = _scatter(model, X, b)
            _, info = torch.linalg.cholesky_ex(_hier_w_inv(model, sigma_bar, S, s, e, eye))
            out.append(info != 0)
    return torch.cat(out)


def prior_scale_objective(model, X, Y, log_scale=0.0, grad=True):
    """Held-out objective of a fitted LocalKNearestKDE as a function of a
    multiplicative scale s = exp(log_scale) of its kNN prior covariances,
    J(s) = mean_m log p(y_m | s): the posteriors are those of
    lakde/local_adaptive_kde.py:83-96 with W_inv_n = s nu0 E[tau_n]
    Sigma0j_n + S_n (responsibilities, nu and the persistent jitters held at
    the fitted values).  Returns (J, dJ/dlog s) (the derivative None when
    grad is False).

    The gradient is the HIP backward of the predictive density (w.r.t. the
    factors L_n^{-1}), chained through each posterior factorisation with the
    float64 Cholesky VJP and contracted with the prior term of W_inv, block by
    block, so no N x D x D float64 tensor is ever formed.  The model's b,
    jitters and records are restored afterwards."""
    require_gpu(X, Y)
    if not hasattr(model, "sigma_0_jitter"):
        raise TypeError("prior_scale_objective needs a LocalKNearestKDE")
    s = float(torch.exp(torch.tensor(float(log_scale), dtype=torch.float64)))
    saved = (model.b.clone(), model.W_inv_jitter.clone(), model.sigma_0_jitter.clone())
    try:
        # s multiplies the prior term nu0 a/b Sigma0j: the posterior kernels
        # see it as b / s (b_out is not written by this recompute)
        model.b.div_(s)
        Linv, nu = posterior_factors(model, X)
        jit = model.W_inv_jitter.clone()
        if not grad:
            with torch.no_grad():
                return float(mixture_log_density(Y, Linv, X, nu).mean()), None
        Linv.requires_grad_(True)
        J = mixture_log_density(Y, Linv, X, nu).mean()
        J.backward()
        gLinv = Linv.grad
        N, D = X.shape
        g = torch.zeros((), dtype=torch.float64, device=X.device)
        for b in model._blocks(N):
            st, e, nb = model._span(b, N)
            S_b, _ = _scatter(model, X, b)
            gA, L = _factor_inverse_vjp(Linv[st:e].detach().double(), gLinv[st:e].double())
            W_j = L @ L.mT                                   # W_inv + jitter I
            prior_term = W_j - S_b.double() - torch.diag_embed(
                jit[st:e].double()[:, None].expand(nb, D))   # s nu0 E[tau] Sigma0j
            g += (gA * prior_term).sum()
        return float(J.detach()), float(g)
    finally:
        model.b.copy_(saved[0])
        model.W_inv_jitter.copy_(saved[1])
        model.sigma_0_jitter.copy_(saved[2])
        _recompute_records(model, X)  # records of the fitted model again


def _recompute_records(model, X):
    N = X.size(0)
    for b in model._blocks(N):
        s, e, nb = model._span(b, N)
        model._block_posterior(X, b, model._nu_s[:nb], model._logdet_s[:nb], model._trace_s[:nb],
                               model._logdet_prior_s[:nb] if hasattr(model, "sigma_0_jitter")
                               else None, update_b=False)
    model._sync(None)


def posterior_factors(model, X):
    """(Linv (N x D x D), nu (N)) of every kernel from the HIP posterior kernel
    (the persistent jitters are updated as by log_pred_density)."""
    N, D = X.shape
    Linv = torch.empty(N, D, D, device=X.device)
    nu = torch.empty(N, device=X.device)
    for b in model._blocks(N):
        s, e, nb = model._span(b, N)
        model._block_posterior(X, b, model._nu_s[:nb], model._logdet_s[:nb], model._trace_s[:nb],
                               model._logdet_prior_s[:nb] if hasattr(model, "sigma_0_jitter")
                               else None, update_b=False)
        Linv[s:e] = model._Linv[:nb]
        nu[s:e] = model._nu_s[:nb]
    model._sync(None)
    return Linv, nu


__all__ = ["MixtureLogDensity", "FactorInverse", "mixture_log_density", "log_pred_density",
           "hierarchical_factors", "unfactorable", "posterior_factors",
           "prior_scale_objective"]
