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


def column_order(X, chunk=64):
    """Column order for the pair-skipping E-step: the points sorted by their
    Morton (Z-order) code on a 2^b grid per coordinate (b = min(10, 62 // D)),
    so that every 64 consecutive columns form a spatially compact chunk, and
    each chunk's bounding box.  Returns (col_perm int32 [N], chunk_box
    float32 [ceil(N/64), 2, D]: lo then hi).  Setup glue (one sort), run
    once per fit."""
    N, D = X.shape
    bits = max(1, min(10, 62 // D))
    lo = X.amin(0)
    span = (X.amax(0) - lo).clamp_min(1e-30)
    q = ((X - lo) / span * (2 ** bits - 1)).round().long().clamp_(0, 2 ** bits - 1)
    code = torch.zeros(N, dtype=torch.long, device=X.device)
    for b in range(bits - 1, -1, -1):
        for d in range(D):
            code = (code << 1) | ((q[:, d] >> b) & 1)
    perm = torch.argsort(code, stable=True)
    nch = (N + chunk - 1) // chunk
    Xp = X[perm]
    if nch * chunk > N:  # pad the last chunk with its own last point
        Xp = torch.cat([Xp, Xp[-1:].expand(nch * chunk - N, D)])
    Xp = Xp.view(nch, chunk, D)
    box = torch.stack([Xp.amin(1), Xp.amax(1)], dim=1).contiguous()
    return perm.to(torch.int32).contiguous(), box.to(torch.float32)
