"""The reference's CPU compute path, for the CPU baseline.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``lakde_amd``) may import
this module; only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py``
use it.

With LAKDE_NO_EXT=1 the reference runs its two hot ops as multi-threaded
torch CPU code (lakde/kernels.py:20-40): ``chol_quad_form`` as one batched
matmul over the difference tensor, ``calc_W_inv`` as a per-row loop of small
matmuls.  The numpy oracle (``oracle.lakde_np``) states the same mathematics
with einsum, single threaded: at D = 63 it runs some forty times slower than
the reference's CPU path on the same host.  ``reference_cpu_ops`` swaps torch
restatements of the two ops into ``oracle.lakde_np`` for the duration of a
block, so that the oracle's model code (the reference's algorithm, line by
line) runs at the speed of the reference's CPU path.  Both restatements are
checked against the numpy ops in tests/test_cpu_path.py.
"""

import contextlib

import numpy as np
import torch

from . import lakde_np as O

__all__ = ["chol_quad_form", "calc_W_inv", "reference_cpu_ops"]


def _tdt(dt):
    return torch.float64 if np.dtype(dt) == np.float64 else torch.float32


def chol_quad_form(W_triu, Xn, Xm, dt=None, chunk=4096):
    """Q[n, m] = ||W_triu[n]^T (Xm[m] - Xn[n])||^2 as the reference's CPU
    path computes it (lakde/kernels.py:20-23): diff = Xm - Xn, LX = W^T
    diff^T (batched matmul), Q = sum of LX^2 over the rows; column chunks of
    ``chunk`` bound the difference tensor."""
    dt = dt or W_triu.dtype
    tdt = _tdt(dt)
    W = torch.from_numpy(np.ascontiguousarray(W_triu, dt)).to(tdt)
    xn = torch.from_numpy(np.ascontiguousarray(Xn, dt)).to(tdt)
    xm = torch.from_numpy(np.ascontiguousarray(Xm, dt)).to(tdt)
    Wt = W.mT
    out = np.empty((xn.shape[0], xm.shape[0]), dt)
    for s in range(0, xm.shape[0], chunk):
        diff = xm[None, s:s + chunk] - xn[:, None]
        LX = Wt @ diff.mT
        out[:, s:s + chunk] = LX.pow_(2).sum(dim=-2).numpy()
    return out


def calc_W_inv(rowptr, col, value, Xn, Xm, dt=None):
    """S_n = sum_e value[e] (Xm[col[e]] - Xn[n]) (...)^T as the reference's
    CPU path computes it (lakde/kernels.py:26-39): one (rnm diff)^T diff
    matmul per row."""
    dt = dt or value.dtype
    tdt = _tdt(dt)
    xn = torch.from_numpy(np.ascontiguousarray(Xn, dt)).to(tdt)
    xm = torch.from_numpy(np.ascontiguousarray(Xm, dt)).to(tdt)
    v_all = torch.from_numpy(np.ascontiguousarray(value, dt)).to(tdt)
    c_all = torch.from_numpy(np.ascontiguousarray(col, np.int64))
    rp = np.asarray(rowptr, np.int64)
    N, D = xn.shape
    out = torch.empty(N, D, D, dtype=tdt)
    for i in range(N):
        a, b = int(rp[i]), int(rp[i + 1])
        diff = xm[c_all[a:b]] - xn[i]
        out[i] = (v_all[a:b, None] * diff).mT @ diff
    return out.numpy()


@contextlib.contextmanager
def reference_cpu_ops(threads):
    """Run the oracle's model code with the torch CPU ops above on
    ``threads`` threads; restores the numpy ops and torch's thread count on
    exit."""
    saved = (O.chol_quad_form, O.calc_W_inv, torch.get_num_threads())
    O.chol_quad_form, O.calc_W_inv = chol_quad_form, calc_W_inv
    torch.set_num_threads(int(threads))
    try:
        yield
    finally:
        O.chol_quad_form, O.calc_W_inv = saved[0], saved[1]
        torch.set_num_threads(saved[2])
