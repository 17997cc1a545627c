"""D x D linear algebra for the shared-prior update (device-side torch glue;
the per-point factorisations run in the HIP posterior kernel).

Restates lakde/linalg_util.py:49-52 (tril_inverse) and :109-170
(stabilize_/stabilized_cholesky) for single matrices and batches.
"""

import torch


def tril_inverse(L):
    D = L.size(-1)
    eye = torch.eye(D, dtype=L.dtype, device=L.device).expand_as(L)
    return torch.linalg.solve_triangular(L, eye, upper=False)


def _factor(X, jitter):
    A = X.clone()
    A.diagonal(dim1=-2, dim2=-1).add_(jitter[..., None] if jitter.dim() else jitter)
    return torch.linalg.cholesky_ex(A, check_errors=False)


def _stabilize(X):
    """Cumulative jitter eps * 2^k (k < 32) until every matrix factors."""
    eps = torch.finfo(X.dtype).eps
    jitter = X.new_zeros(X.shape[:-2])
    mask = X.new_ones(X.shape[:-2], dtype=torch.bool)
    for k in range(32):
        jitter[mask] += eps * 2 ** k
        _, info = _factor(X, jitter)
        mask = info != 0
        if not torch.any(mask):
            return jitter
    raise RuntimeError("Stabilization did not converge")


def stabilized_cholesky(X, jitter=None):
    """Cholesky with a persistent jitter; returns (L, jitter).  The jitter
    keeps its own dtype (the reference stores the shared prior's jitter as a
    0-dim float32 tensor while factoring in float64)."""
    if jitter is not None and torch.count_nonzero(jitter) != 0:
        L, info = _factor(X, jitter.to(X.dtype))
    else:
        L, info = torch.linalg.cholesky_ex(X, check_errors=False)
    if not torch.any(info):
        return L, jitter
    if X.dim() == 2:
        new = _stabilize(X[None])[0]
        jitter = (X.new_zeros(()) if jitter is None else jitter.clone())
        jitter.fill_(new.to(jitter.dtype))
    else:
        jitter = X.new_zeros(X.shape[:-2]) if jitter is None else jitter.clone()
        mask = info != 0
        jitter[mask] = _stabilize(X[mask]).to(jitter.dtype)
    return stabilized_cholesky(X, jitter)
