"""GPU parity of the kNN graph kernel (lakde_knn_graph; reference
lakde/utils.py:77-116) against the reference's golden graph, the oracle
(float64 exact distances, equal distances ordered by index) and brute force.

Bar: bit-exact indices wherever the order is decided by more than FP32
rounding.  Two distinct points whose float64 distances to a query differ by
less than 1e-5 relative may swap (the kernel ranks FP32 distances); every
such swap is checked to be one.
"""

import numpy as np
import pytest
import torch

from oracle import lakde_np as O

pytestmark = pytest.mark.gpu

DEV = "cuda"
REL = 1e-5


def _knn(X, k, include_self=False, dtype=torch.int32, return_dists=False):
    from lakde_amd import _fused
    Xt = torch.as_tensor(np.ascontiguousarray(X, np.float32), device=DEV)
    res = _fused.knn_graph(Xt, k, include_self=include_self, dtype=dtype,
                           return_dists=return_dists)
    torch.cuda.synchronize()
    if return_dists:
        return res[0].cpu().numpy(), res[1].cpu().numpy()
    return res.cpu().numpy()


def _d64(X, rows, idx):
    X64 = np.asarray(X, np.float64)
    return np.sum((X64[idx] - X64[rows][:, None]) ** 2, axis=-1)


def _check_against(X, got, want, rows=None):
    """Equal, or different only where float64 distances tie within REL."""
    rows = np.arange(got.shape[0]) if rows is None else rows
    assert got.shape == want.shape
    bad = np.nonzero((got != want).any(axis=1))[0]
    if bad.size == 0:
        return 0
    dg = _d64(X, rows[bad], got[bad])
    dw = _d64(X, rows[bad], want[bad])
    # the same distance profile, position by position, up to FP32 rounding
    np.testing.assert_allclose(dg, dw, rtol=REL, atol=1e-12)
    # and every row of the kernel is sorted by float64 distance up to rounding
    assert np.all(np.diff(dg, axis=1) >= -REL * np.abs(dg[:, 1:]) - 1e-12)
    return bad.size


def test_knn_graph_golden(golden):
    g = golden("ops")
    got = _knn(g["knn_X"], 6)
    np.testing.assert_array_equal(got, g["knn_graph"])


@pytest.mark.parametrize("N,D,k,include_self", [
    (300, 1, 5, False),
    (8000, 3, 8, False),      # 16 candidate splits merged per query
    (3000, 2, 20, False),     # C1-like: D=2, k=20
    (5000, 6, 64, False),     # C2 dimension and k
    (2000, 21, 128, True),    # C3 dimension and k, hierarchical (self included)
    (1500, 43, 256, False),   # C4 dimension and k
    (1200, 63, 128, False),   # C5 dimension and k
])
def test_knn_graph_matches_oracle(N, D, k, include_self):
    rng = np.random.default_rng(N + D)
    X = (rng.standard_normal((N, D)) * rng.uniform(0.5, 2.0, D)).astype(np.float32)
    print(f"knn N={N} D={D} k={k} include_self={include_self}", flush=True)
    got = _knn(X, k, include_self=include_self)
    want = O.compute_knn_graph(X, k, include_self=include_self)
    nbad = _check_against(X, got, want)
    assert nbad <= max(2, N // 100)
    if not include_self:
        assert not (got == np.arange(N)[:, None]).any()


def test_knn_graph_exact_ties_by_index():
    # 40 distinct points, each repeated 9 times: every query has exact ties
    rng = np.random.default_rng(3)
    base = rng.standard_normal((40, 5)).astype(np.float32)
    X = np.repeat(base, 9, axis=0)
    perm = rng.permutation(X.shape[0])
    X = X[perm]
    for include_self in (False, True):
        for k in (1, 8, 9, 30):
            got = _knn(X, k, include_self=include_self, dtype=torch.int64)
            want = O.compute_knn_graph(X, k, include_self=include_self)
            np.testing.assert_array_equal(got, want)


def test_knn_graph_dists_and_int64():
    rng = np.random.default_rng(11)
    X = rng.standard_normal((700, 9)).astype(np.float32)
    d, idx = _knn(X, 17, dtype=torch.int64, return_dists=True)
    assert idx.dtype == np.int64
    want = O.compute_knn_graph(X, 17)
    _check_against(X, idx, want)
    np.testing.assert_allclose(d, np.sqrt(_d64(X, np.arange(700), idx)), rtol=1e-5, atol=1e-6)


def test_knn_graph_single_split_large():
    # N large enough for one candidate split per query (>= 2048 blocks of 256
    # queries); checked against float64 brute force on a sample of queries
    N, D, k = 600_000, 3, 12
    rng = np.random.default_rng(7)
    X = rng.standard_normal((N, D)).astype(np.float32)
    got = _knn(X, k)
    rows = rng.choice(N, 64, replace=False)
    X64 = X.astype(np.float64)
    want = np.empty((rows.size, k), np.int64)
    for t, r in enumerate(rows):
        d = np.sum((X64 - X64[r]) ** 2, axis=1)
        d[r] = np.inf
        cand = np.argpartition(d, k)[:k + 1]
        order = np.lexsort((cand, d[cand]))[:k]
        want[t] = cand[order]
    nbad = _check_against(X, got[rows].astype(np.int64), want, rows=rows)
    assert nbad <= 2
    assert not (got == np.arange(N)[:, None]).any()


def test_knn_graph_rejects_bad_k():
    from lakde_amd import _fused
    from lakde_amd._lib import LakdeError
    X = torch.zeros(10, 2, device=DEV)
    with pytest.raises(LakdeError):
        _fused.knn_graph(X, 10, include_self=False)
    _fused.knn_graph(X, 10, include_self=True)
    torch.cuda.synchronize()
