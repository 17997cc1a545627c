"""The CPU baseline's torch restatement of the reference's CPU ops
(oracle/cpu_path.py) against the numpy oracle ops, which the golden vectors
pin, and the oracle's model with those ops swapped in against the golden
fit of the reference."""

import numpy as np
import pytest

from oracle import cpu_path as C
from oracle import lakde_np as O


@pytest.mark.parametrize("D,dt", [(2, np.float32), (21, np.float64), (63, np.float32)])
def test_quad_form_matches_numpy(D, dt):
    rng = np.random.default_rng(D)
    W = np.triu(rng.standard_normal((7, D, D))).astype(dt)
    Xn = rng.standard_normal((7, D)).astype(dt)
    Xm = rng.standard_normal((300, D)).astype(dt)
    a = C.chol_quad_form(W, Xn, Xm, dt=dt, chunk=128)
    b = O.chol_quad_form(W, Xn, Xm, dt=dt)
    assert a.dtype == b.dtype
    np.testing.assert_allclose(a, b, rtol=1e-5 if dt == np.float32 else 1e-12)


@pytest.mark.parametrize("D,dt", [(3, np.float32), (17, np.float64)])
def test_w_inv_matches_numpy(D, dt):
    rng = np.random.default_rng(D)
    N, M = 9, 40
    counts = rng.integers(0, 6, N)
    counts[2] = 0
    rowptr = np.concatenate([[0], np.cumsum(counts)])
    col = rng.integers(0, M, rowptr[-1]).astype(np.int32)
    val = rng.random(rowptr[-1]).astype(dt)
    Xn = rng.standard_normal((N, D)).astype(dt)
    Xm = rng.standard_normal((M, D)).astype(dt)
    a = C.calc_W_inv(rowptr, col, val, Xn, Xm, dt=dt)
    b = O.calc_W_inv(rowptr, col, val, Xn, Xm, dt=dt)
    np.testing.assert_allclose(a, b, rtol=1e-5 if dt == np.float32 else 1e-12, atol=1e-6)
    assert np.all(a[2] == 0)


def test_swapped_oracle_fit_matches_reference_golden():
    """The oracle's fit run with the swapped ops reproduces the reference's
    held-out log density on the small circle fixture (1e-4 relative)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "knn_circle8_small.npz"))
    thr = float(g["threshold"]) or None
    with C.reference_cpu_ops(2):
        m = O.LocalKNearestKDE(k_or_knn_graph=g["graph"], block_size=int(g["block_size"]))
        m.fit(g["X"], iterations=int(g["iterations"]), sparsity_threshold=thr)
        lp = m.log_pred_density(g["X"], g["Y"])
    assert O.chol_quad_form is not C.chol_quad_form  # restored
    ref = g["log_pred"]
    assert abs(lp.mean() - ref.mean()) <= 1e-4 * abs(ref.mean())
