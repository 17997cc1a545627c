"""Initialisation helpers (device-side torch glue, init only) and the SIGINT
deferral around model state updates."""

import signal
import threading

import torch


class InterruptDelay:
    """Defer SIGINT while a model's device and host state are updated together
    (the semantics of lakde/utils.py:240-255): a SIGINT that arrives inside
    the block is recorded, and when the block exits the previous handler
    receives it (the default handler raises KeyboardInterrupt there, after
    the bookkeeping).  Outside the main thread Python delivers no signals to
    handlers, so the block is a no-op there."""

    def __init__(self):
        self.signal_received = None
        self.old_handler = None

    def __enter__(self):
        self.signal_received = None
        self.old_handler = None
        if threading.current_thread() is threading.main_thread():
            self.old_handler = signal.signal(signal.SIGINT, self._handler)
        return self

    def _handler(self, sig, frame):
        self.signal_received = (sig, frame)

    def __exit__(self, exc_type, exc, tb):
        if self.old_handler is None:
            return False
        signal.signal(signal.SIGINT, self.old_handler)
        got = self.signal_received
        self.signal_received = None
        if got is not None:
            if callable(self.old_handler):
                self.old_handler(*got)
            elif self.old_handler == signal.SIG_DFL:
                signal.raise_signal(signal.SIGINT)
        return False


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
    """Column order for the pair-skipping E-step: a balanced k-d ordering of
    the points whose leaves are the 64-column chunks (each node is split on
    its widest coordinate at half its chunks, so the splits fall on chunk
    boundaries), and each chunk's bounding box.  Returns (col_perm int32
    [N], chunk_box float32 [ceil(N/64), 2, D]: lo then hi).  Setup glue (two
    sorts per tree level), run once per fit."""
    N, D = X.shape
    dev = X.device
    nch = (N + chunk - 1) // chunk
    perm = torch.arange(N, device=dev)
    pos = torch.arange(N, device=dev)
    start = torch.zeros(N, dtype=torch.long, device=dev)  # node = its first chunk
    cnt = torch.full((N,), nch, dtype=torch.long, device=dev)  # the node's chunks
    while bool((cnt > 1).any()):
        active = cnt > 1
        Xp = X[perm]
        idx = start[:, None].expand(N, D)
        hi = torch.full((nch, D), -float("inf"), device=dev).scatter_reduce(0, idx, Xp, "amax")
        lo = torch.full((nch, D), float("inf"), device=dev).scatter_reduce(0, idx, Xp, "amin")
        dim = (hi - lo).argmax(1)
        val = Xp.gather(1, dim[start][:, None]).squeeze(1)
        # within each node by the split coordinate (nodes stay contiguous)
        o = torch.argsort(val, stable=True)
        o = o[torch.argsort(start[o], stable=True)]
        perm = perm[o]
        left = cnt // 2
        right = (pos - start * chunk) >= left * chunk
        start, cnt = (torch.where(active & right, start + left, start),
                      torch.where(active, torch.where(right, cnt - left, left), cnt))
    Xp = X[perm]
    if nch * chunk > N:  # pad the last chunk with its own last point
        Xp = torch.cat([Xp, Xp[-1:].expand(nch * chunk - N, D)])
    Xp = Xp.view(nch, chunk, D)
    box = torch.stack([Xp.amin(1), Xp.amax(1)], dim=1).contiguous()
    return perm.to(torch.int32).contiguous(), box.to(torch.float32)
