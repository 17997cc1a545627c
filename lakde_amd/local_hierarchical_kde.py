"""LocalHierarchicalKDE on the fused HIP path
(mirrors lakde/local_hierarchical_kde.py:77-495).

The per-point work (posterior factorisation, E-step, rebalance) runs on the
HIP kernels; the D x D shared-prior update (FP64 sums, one stabilised
Cholesky and triangular inverse per partial step) is device-side torch glue.
"""

import math
from typing import Union

import numpy as np
import torch

from ._lib import PRIOR_SHARED
from .linalg_util import stabilized_cholesky, tril_inverse
from .local_adaptive_kde import LocallyAdaptiveKDE
from .utils import InterruptDelay, compute_knn_graph, cov, tensor_as


class LocalHierarchicalKDE(LocallyAdaptiveKDE):
    def __init__(self, nu_0: Union[torch.Tensor, float] = None,
                 nu_shared_0: Union[torch.Tensor, float] = None,
                 k_or_knn_graph: Union[torch.Tensor, int] = None, block_size: int = 512,
                 a_0: Union[torch.Tensor, float] = 1e-6, b_0: Union[torch.Tensor, float] = 1e-6,
                 verbose=False, logs=False, precision="fp32", rebalance="lazy",
                 estep_skip=True):
        super().__init__(nu_0, a_0, b_0, block_size, verbose, logs, precision, rebalance,
                         estep_skip)
        self.k_or_knn_graph = k_or_knn_graph
        self.k = None
        self.nu_shared_0 = nu_shared_0
        self.lambda_0_triu = None
        self.lambda_0 = None
        self.nu_shared = None
        self.W_shared_triu = None
        self.expect_lambda_contribs = None
        self.W_shared_inv_jitter = None

    # lakde/local_hierarchical_kde.py:111-196
    def init_parameters_(self, X, sparsity_threshold=None):
        self._common_init(X, sparsity_threshold)
        N, D = X.shape
        self.W_shared_inv_jitter = X.new_zeros(())
        if self.nu_shared_0 is not None:
            self.nu_shared_0 = tensor_as(self.nu_shared_0, X)
        else:
            self.nu_shared_0 = X.new_full((), D - 1 + 1e-2)
        if self.nu_shared_0 <= D - 1:
            raise ValueError("nu_shared_0 must be greater than the data dimensionality minus one, "
                             "or the ELBO will be undefined")
        sigma_0 = cov(X, dtype=torch.double)
        sigma_0_tril, _ = stabilized_cholesky(sigma_0)
        self.lambda_0_triu = tril_inverse(sigma_0_tril).T
        self.lambda_0 = self.lambda_0_triu @ self.lambda_0_triu.T
        g = self.k_or_knn_graph
        if isinstance(g, (int, np.integer)):
            g = compute_knn_graph(X, int(g), 1024, include_self=True, dtype=torch.int32)
        elif isinstance(g, np.ndarray):
            g = torch.from_numpy(g)
        elif not (torch.is_tensor(g) and g.dtype in (torch.int32, torch.int64)):
            raise ValueError("k_or_knn_graph type not recognised")
        graph = g.to(X.device).contiguous()
        del self.k_or_knn_graph
        self.k = graph.size(-1)
        self._init_arena(X, graph, sparsity_threshold)
        self.nu_shared = self.nu_shared_0 + N * self.nu_0.double()
        self.W_shared_triu = self.nu_0 * sigma_0_tril.T / self.nu_shared
        self.nu = X.new_empty(N)
        self.a = (self.a_0 + self.nu_0 * D / 2).expand(N).contiguous()
        self.b = self.a.clone()
        self.rnm_log_consts = X.new_full((N,), -math.inf)
        self.logdet_W = X.new_empty(N)
        num_blocks = (N - 1) // self.block_size + 1
        self.expect_lambda_contribs = X.new_empty(num_blocks, D, D, dtype=torch.double)
        self._alloc(X)

    # lakde/local_hierarchical_kde.py:269-271
    def calc_prior_cov(self, *_):
        t = torch.sqrt(self.nu_shared) * self.W_shared_triu
        return t @ t.T, t

    # lakde/local_hierarchical_kde.py:260-267
    def calc_prior_prec(self):
        return (self.expect_lambda_contribs.sum(dim=0).mul_(self.nu_0)
                .add_(self.lambda_0, alpha=float(self.nu_shared_0)))

    def _prior_spec(self, X, b, nb):
        return PRIOR_SHARED, self._sigma_bar32, None, None, None

    def _step_outputs(self, b, nb):
        s = b * self.block_size
        e = s + nb
        return self.nu[s:e], self.logdet_W[s:e], self._trace_s[:nb], None

    def _prior_ptr(self, b, current):
        return self._sigma_bar32.data_ptr()

    def _set_sigma_bar(self):
        sigma_bar, _ = self.calc_prior_cov()
        self._sigma_bar32 = sigma_bar.to(torch.float32).contiguous()

    def _lambda_contrib(self, nb, nu, b_blk, s, e):
        """sum_n (U_n U_n^T) nu_n E[tau_n] in float32 per point, summed in
        float64 (lakde/local_hierarchical_kde.py:214-219, 293-297)."""
        Linv = self._Linv[:nb]
        UUt = torch.bmm(Linv.mT.contiguous(), Linv)
        UUt *= (nu * (self.a[s:e] / b_blk))[:, None, None]
        return torch.sum(UUt, dtype=torch.double, dim=0)

    # lakde/local_hierarchical_kde.py:198-258
    def init_summaries_(self, X, sparsity_threshold=None):
        N = X.size(0)
        self._set_sigma_bar()
        for b in self._blocks(N):
            s, e, nb = self._span(b, N)
            self._block_posterior(X, b, self.nu[s:e], self.logdet_W[s:e], self._trace_s[:nb],
                                  None, update_b=False)
            self.expect_lambda_contribs[b] = self._lambda_contrib(nb, self.nu[s:e], self.b[s:e],
                                                                  s, e)
            self._init_lse_pass(X, b)
        self._active = 0.0
        for b in self._blocks(N):
            s, e, nb = self._span(b, N)
            self._block_posterior(X, b, self._nu_s[:nb], self._logdet_s[:nb],
                                  self._trace_s[:nb], None, update_b=False)
            self._init_fixed_pass(X, b, sparsity_threshold)
        self._init_rebalance()

    # lakde/local_hierarchical_kde.py:273-311
    def partial_step_(self, X, block_idx, sparsity_threshold=None):
        N = X.size(0)
        s, e, nb = self._span(block_idx, N)
        # the step and the shared-prior update complete together; a SIGINT
        # in between is delivered afterwards (lakde/local_hierarchical_kde.py:299)
        with InterruptDelay():
            self._set_sigma_bar()
            # b[s:e] <- b0 + nu0 * trace / 2 is written by the posterior kernel
            # after E[tau] has been read; nothing in the E-step reads b
            self._partial_step(X, block_idx, sparsity_threshold)
            # nu, b and the factors are not touched by the E-step
            self.expect_lambda_contribs[block_idx] = self._lambda_contrib(
                nb, self.nu[s:e], self.b[s:e], s, e)
            W_shared_inv = self.calc_prior_prec()
            W_shared_inv_tril, self.W_shared_inv_jitter = stabilized_cholesky(
                W_shared_inv, self.W_shared_inv_jitter)
            self.W_shared_triu = tril_inverse(W_shared_inv_tril).T

    def log_pred_density(self, X, Y):
        """lakde/local_adaptive_kde.py:200-223 with calc_prior_cov() of the
        current shared prior factor (lakde/local_hierarchical_kde.py:269-271):
        the posteriors are recomputed from the W_shared_triu of the last
        shared-prior update, not from the prior the last partial step read."""
        if self.W_shared_triu is None:
            raise RuntimeError("fit the model first")
        self._set_sigma_bar()
        return super().log_pred_density(X, Y)

    def state_dict(self):
        d = super().state_dict()
        d.update({"k": self.k, "nu_shared_0": self.nu_shared_0, "lambda_0_triu": self.lambda_0_triu,
                  "lambda_0": self.lambda_0, "nu_shared": self.nu_shared,
                  "W_shared_triu": self.W_shared_triu,
                  "expect_lambda_contribs": self.expect_lambda_contribs,
                  "W_shared_inv_jitter": self.W_shared_inv_jitter})
        return d

    def load_state_dict(self, state_dict, X):
        """lakde/local_hierarchical_kde.py:463-490; X sizes the device buffers."""
        self._load_common(state_dict, X.device)
        for k in ("k", "nu_shared_0", "lambda_0_triu", "lambda_0", "nu_shared", "W_shared_triu",
                  "expect_lambda_contribs", "W_shared_inv_jitter"):
            v = state_dict[k]
            setattr(self, k, v.to(X.device).clone() if torch.is_tensor(v) else v)
        self._load_blocks(list(state_dict["rnm_blocks"]), X)
        self._finish_load(X)
