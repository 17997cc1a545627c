"""Seeded synthetic workloads for the five fit/predict configurations.

Every generator draws from ``numpy.random.default_rng(seed)`` and returns
``(X_train, X_test)`` as contiguous float32 arrays, so CPU and GPU runs see the
same bytes.  The shapes and distributions are:

* ``circle8``    -- N=5,000 / 1,000, D=2, 8 isotropic Gaussians (std 0.5) with
  means evenly spaced on a circle of radius 4, equal weights (seed 0).
* ``mix32_d6``   -- N=200,000 / 20,000, D=6, 32 Gaussians with covariance
  A A^T + 0.01 I, A ~ N(0, 0.3^2); coordinate 0 replaced by U[0, 1) noise and
  coordinate 1 quantised to 24 levels (seed 1).  The quantised column makes
  many kNN covariances singular, which exercises the jitter path.
* ``mix64_d21``  -- N=300,000 / 30,000, D=21, 64 Gaussians whose covariance
  eigenvalues are log-uniform in [1e-3, 1] in random orthogonal bases, plus 5%
  Student-t (nu=3) background (seed 2).
* ``latent_d43`` -- N=30,000 / 3,000, D=43, x = z W^T + eps with z ~ N(0, I_10),
  W ~ N(0, 1/10), eps ~ N(0, 0.05^2 I); two columns replaced by integers in
  {0..7} (seed 3).
* ``mix128_d63`` -- N=1,000,000 / 50,000, D=63, 128 Gaussians with covariance
  A A^T + 1e-3 I, A ~ N(0, 0.2^2), rescaled to unit marginal variance (seed 4).

``n_train``/``n_test`` overrides draw a prefix-compatible smaller sample of the
same distribution (used by the parity tests and the CPU baseline sample).
"""

import hashlib

import numpy as np

__all__ = ["WORKLOADS", "make", "digest"]


def _gmm(rng, n, means, chols):
    k, d = means.shape
    comp = rng.integers(0, k, size=n)
    z = rng.standard_normal((n, d))
    x = np.empty((n, d))
    for c in range(k):
        sel = comp == c
        x[sel] = means[c] + z[sel] @ chols[c].T
    return x


def _circle8(rng, n):
    ang = 2 * np.pi * np.arange(8) / 8
    means = 4.0 * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    comp = rng.integers(0, 8, size=n)
    return means[comp] + 0.5 * rng.standard_normal((n, 2))


def _mix32_d6(rng, n):
    d, k = 6, 32
    means = rng.normal(0.0, 1.5, size=(k, d))
    a = rng.normal(0.0, 0.3, size=(k, d, d))
    cov = a @ np.swapaxes(a, 1, 2) + 0.01 * np.eye(d)
    x = _gmm(rng, n, means, np.linalg.cholesky(cov))
    x[:, 0] = rng.random(n)
    lo, hi = x[:, 1].min(), x[:, 1].max()
    lev = np.clip(np.floor((x[:, 1] - lo) / (hi - lo) * 24), 0, 23)
    x[:, 1] = lo + (lev + 0.5) * (hi - lo) / 24
    return x


def _mix64_d21(rng, n):
    d, k = 21, 64
    means = rng.normal(0.0, 2.0, size=(k, d))
    chols = np.empty((k, d, d))
    for c in range(k):
        q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        ev = np.exp(rng.uniform(np.log(1e-3), 0.0, size=d))
        chols[c] = np.linalg.cholesky((q * ev) @ q.T + 1e-12 * np.eye(d))
    x = _gmm(rng, n, means, chols)
    bg = rng.random(n) < 0.05
    nb = int(bg.sum())
    g = rng.standard_normal((nb, d))
    u = rng.chisquare(3.0, size=(nb, 1))
    x[bg] = 2.0 * g / np.sqrt(u / 3.0)
    return x


def _latent_d43(rng, n):
    d, q = 43, 10
    w = rng.normal(0.0, np.sqrt(0.1), size=(d, q))
    z = rng.standard_normal((n, q))
    x = z @ w.T + rng.normal(0.0, 0.05, size=(n, d))
    x[:, 3] = rng.integers(0, 8, size=n)
    x[:, 17] = rng.integers(0, 8, size=n)
    return x


def _mix128_d63(rng, n):
    d, k = 63, 128
    means = rng.normal(0.0, 1.0, size=(k, d))
    a = rng.normal(0.0, 0.2, size=(k, d, d))
    cov = a @ np.swapaxes(a, 1, 2) + 1e-3 * np.eye(d)
    x = _gmm(rng, n, means, np.linalg.cholesky(cov))
    return x


# name -> (generator, seed, N_train, N_test, D, k, model, threshold, iterations)
WORKLOADS = {
    "circle8": dict(gen=_circle8, seed=0, n=5000, m=1000, d=2, k=20,
                    model="knn", threshold=0.0, iterations=50),
    "mix32_d6": dict(gen=_mix32_d6, seed=1, n=200_000, m=20_000, d=6, k=64,
                     model="knn", threshold=1 / 64, iterations=30),
    "mix64_d21": dict(gen=_mix64_d21, seed=2, n=300_000, m=30_000, d=21, k=128,
                      model="knn", threshold=1 / 128, iterations=30),
    "latent_d43": dict(gen=_latent_d43, seed=3, n=30_000, m=3_000, d=43, k=256,
                       model="hierarchical", threshold=1 / 256, iterations=30),
    "mix128_d63": dict(gen=_mix128_d63, seed=4, n=1_000_000, m=50_000, d=63,
                       k=128, model="knn", threshold=1 / 128, iterations=20),
}


def make(name, n_train=None, n_test=None):
    """Return (X_train, X_test) float32 arrays for workload ``name``.

    The full sample (N + M rows) is always drawn, then the first ``n_train``
    rows are the training set and the next ``n_test`` rows the test set, so a
    reduced sample is a subset of the same draw.  For ``mix128_d63`` columns
    are standardised over the whole draw before the split.
    """
    w = WORKLOADS[name]
    rng = np.random.default_rng(w["seed"])
    n, m = w["n"], w["m"]
    x = w["gen"](rng, n + m)
    if name == "mix128_d63":
        x = (x - x.mean(axis=0)) / x.std(axis=0)
    n_train = n if n_train is None else n_train
    n_test = m if n_test is None else n_test
    if n_train > n or n_test > m:
        raise ValueError("requested sample larger than the workload")
    tr = np.ascontiguousarray(x[:n_train], dtype=np.float32)
    te = np.ascontiguousarray(x[n:n + n_test], dtype=np.float32)
    return tr, te


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]
