"""Initialisation helpers (device-side torch glue, init only)."""

import torch


def tensor_as(x, other):
    """lakde/utils.py:8-11."""
    if torch.is_tensor(x):
        return x.to(dtype=other.dtype, device=other.device)
    return torch.tensor(x, dtype=other.dtype, device=other.device)


def cov(X, dtype=torch.float64):
    """Sample covariance (lakde/utils.py:28-37)."""
    N = X.size(-2)
    X_bar = X.mean(dim=-2, dtype=dtype, keepdim=True)
    Xc = X.to(dtype) - X_bar
    return (Xc.mT @ Xc) / (N - 1)


def compute_knn_graph(X, k, bsize=1024, include_self=False, dtype=torch.int32):
    """k nearest neighbours by L2 distance (lakde/utils.py:77-116, the
    generic path), on the HIP kernel: rows sorted by distance, equal
    distances by index.  Used only at initialisation when the model is given
    an integer k.  ``bsize`` is accepted for signature compatibility; the
    kernel tiles by itself."""
    from . import _fused
    return _fused.knn_graph(X.to(torch.float32), int(k), include_self=include_self, dtype=dtype)
