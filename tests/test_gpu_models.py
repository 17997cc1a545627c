"""End-to-end parity of the fused GPU models against the reference's own
outputs (golden vectors written by running lakde with LAKDE_NO_EXT=1) and the
CPU oracle.

Bars (float32):
* kNN model: held-out mean log density within 1e-4 relative of the
  reference; per-point within 1e-3 relative; persistent jitters equal;
  responsibility index sets equal up to entries within 1e-4 relative of the
  threshold (a float32 rounding difference can move such an entry across it)
  or, for threshold 0, entries whose value is below 1e-30 (exp underflow).
* hierarchical model (FP32 jitter-chaotic, see the oracle tests): held-out
  mean log density within 2e-3 relative.
"""

import numpy as np
import pytest
import torch

from oracle import lakde_np as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _fit(d, cls, rebalance="lazy"):
    X, Y = _t(d["X"]), _t(d["Y"])
    m = cls(k_or_knn_graph=_t(d["graph"]), block_size=int(d["block_size"]), rebalance=rebalance)
    m.fit(X, iterations=int(d["iterations"]), sparsity_threshold=float(d["threshold"]))
    lp = m.log_pred_density(X, Y).cpu().numpy()
    return m, lp


def _pairs(m):
    rows, cols, vals = [], [], []
    for b, (ind, val) in enumerate(m.rnm_blocks):
        ind = ind.cpu().numpy().astype(np.int64)
        rows.append(ind[0] + b * m.block_size)
        cols.append(ind[1])
        vals.append(val.cpu().numpy())
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def _close_but_flips(a, b, rtol=1e-5, atol=1e-5, frac=0.995, flip_tol=0.05):
    """Arrays equal within (rtol, atol) except a few elements, each within
    flip_tol: an entry whose value sits at the threshold to float32
    rounding can be kept in one run and not in the other (the runs sum in
    different orders), and a dropped entry of value t moves its column's
    normaliser by about log(1 - t)."""
    close = np.isclose(a, b, rtol=rtol, atol=atol)
    assert close.mean() >= frac, (close.mean(), np.abs(a - b).max())
    np.testing.assert_allclose(a, b, rtol=0, atol=flip_tol)


def _check_index_sets(rows, cols, vals, d, threshold):
    N = d["X"].shape[0]
    a = rows * N + cols
    b = d["rnm_row"].astype(np.int64) * N + d["rnm_col"].astype(np.int64)
    only_dev = np.setdiff1d(a, b)
    only_ref = np.setdiff1d(b, a)
    val_dev = dict(zip(a.tolist(), vals.tolist()))
    val_ref = dict(zip(b.tolist(), d["rnm_val"].tolist()))
    for k in only_dev:
        v = val_dev[k]
        assert (abs(v - threshold) <= 1e-4 * max(threshold, 1e-30)) or v < 1e-30, (k, v)
    for k in only_ref:
        v = val_ref[k]
        assert (abs(v - threshold) <= 1e-4 * max(threshold, 1e-30)) or v < 1e-30, (k, v)
    return len(only_dev) + len(only_ref)


@pytest.mark.parametrize("rebalance", ["lazy", "exact"])
@pytest.mark.parametrize("name", ["knn_circle8_small", "knn_mix6_small", "knn_degenerate_small"])
def test_knn_model_matches_reference(golden, name, rebalance):
    from lakde_amd import LocalKNearestKDE
    d = golden(name)
    print(f"[{name}] N={d['X'].shape[0]} D={d['X'].shape[1]} bs={int(d['block_size'])} "
          f"rebalance={rebalance}", flush=True)
    m, lp = _fit(d, LocalKNearestKDE, rebalance)
    ref = d["log_pred"]
    rel = abs(lp.mean() - ref.mean()) / abs(ref.mean())
    assert rel <= 1e-4, rel
    np.testing.assert_allclose(lp, ref, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(m.rnm_log_consts.cpu().numpy(), d["log_consts"], rtol=1e-4,
                               atol=1e-3)
    np.testing.assert_allclose(m.nu.cpu().numpy(), d["nu"], rtol=1e-4)
    np.testing.assert_array_equal(m.W_inv_jitter.cpu().numpy(), d["W_inv_jitter"])
    np.testing.assert_array_equal(m.sigma_0_jitter.cpu().numpy(), d["sigma_0_jitter"])
    rows, cols, vals = _pairs(m)
    # canonical order: rows ascending, columns ascending within a row
    key = rows * d["X"].shape[0] + cols
    assert np.all(np.diff(key) > 0)
    _check_index_sets(rows, cols, vals, d, float(d["threshold"]))


@pytest.mark.parametrize("rebalance", ["lazy", "exact"])
def test_knn_circle_full_matches_reference(golden, rebalance):
    """The 5,000-point, 50-iteration circle workload against the reference's
    own CPU run."""
    from lakde_amd import LocalKNearestKDE
    d = golden("knn_circle8_full")
    m, lp = _fit(d, LocalKNearestKDE, rebalance)
    ref = d["log_pred"]
    rel = abs(lp.mean() - ref.mean()) / abs(ref.mean())
    print(f"circle8 full ({rebalance}): mean log p {lp.mean():.6f} (reference {ref.mean():.6f}), "
          f"rel {rel:.2e}, nnz {m.arena.total_nnz()} (reference {int(d['rnm_nnz'])})")
    assert rel <= 1e-4
    np.testing.assert_allclose(lp, ref, rtol=1e-3, atol=1e-3)
    assert abs(m.arena.total_nnz() - int(d["rnm_nnz"])) <= 1e-4 * int(d["rnm_nnz"])


_FP64_SPREAD = {}


def _fp64_spread(d, name):
    """|FP64 oracle - FP32 reference| / |reference| of the held-out mean: how
    far the hierarchical fit moves under a change of rounding alone."""
    if name not in _FP64_SPREAD:
        mo = O.LocalHierarchicalKDE(k_or_knn_graph=d["graph"], block_size=int(d["block_size"]),
                                    dtype=np.float64)
        X = np.asarray(d["X"], np.float64)
        mo.fit(X, iterations=int(d["iterations"]), sparsity_threshold=float(d["threshold"]))
        lp64 = mo.log_pred_density(X, np.asarray(d["Y"], np.float64)).mean()
        ref = d["log_pred"].mean()
        _FP64_SPREAD[name] = abs(lp64 - ref) / abs(ref)
    return _FP64_SPREAD[name]


# sweeps of each hierarchical fixture through which the FP32 trajectory is
# pinned to the reference (tests/test_oracle_golden.py:HIER_PINS holds the
# same table for the oracle and the reasons): state after sweep i checked for
# i < pinned, -1 = initialisation only; init rtol; held-out mean bar (None:
# max(2e-3, 1.2x the reference's own FP32-vs-FP64 spread))
HIER_PINS = {
    "hier_circle8_nojit": (4, 1e-4, 1e-4),
    "hier_normal4_nojit": (0, 1e-4, 1e-4),
    "hier_smooth_small": (0, 1e-4, None),
    "hier_latent_small": (-1, 5e-3, None),
    "hier_mix6_small": (-1, 5e-3, None),
}


def _check_hier_state_gpu(m, d, it, tag):
    """As the oracle test's _check_hier_state: normalisers and nu within 1e-4
    (1e-4 absolute floor), b within 1e-4, equal jitters, and W_shared_triu
    within 5e-5 in Frobenius norm relative to the reference's.  The factor is
    the inverse Cholesky factor of a sum of 1,000 float32 outer products
    whose off-diagonal entries cancel to a few percent of the diagonal: the
    float64 oracle itself lands 1.0e-5 from the reference's float32 value
    after the second sweep of hier_circle8_nojit, so a bar at 1e-5 would
    test the reference's rounding rather than this implementation."""
    if it < 0:
        lc, nu, jit, wst, b = (d["init_log_consts"], d["init_nu"], d["init_W_inv_jitter"],
                               d["init_W_shared_triu"], None)
    else:
        lc, nu, jit, wst, b = (d["traj_log_consts"][it], d["traj_nu"][it],
                               d["traj_W_inv_jitter"][it], d["traj_W_shared_triu"][it],
                               d["traj_b"][it])
    np.testing.assert_allclose(m.rnm_log_consts.cpu().numpy(), lc, rtol=1e-4, atol=1e-4,
                               err_msg=tag)
    np.testing.assert_allclose(m.nu.cpu().numpy(), nu, rtol=1e-4, err_msg=tag)
    if b is not None:
        np.testing.assert_allclose(m.b.cpu().numpy(), b, rtol=1e-4, err_msg=tag)
    np.testing.assert_array_equal(m.W_inv_jitter.cpu().numpy(), jit, err_msg=tag)
    wd = m.W_shared_triu.cpu().numpy()
    wrel = np.linalg.norm(wd - wst) / np.linalg.norm(wst)
    lrel = np.max(np.abs(m.rnm_log_consts.cpu().numpy() - lc) / np.maximum(np.abs(lc), 1.0))
    print(f"{tag}: normalisers {lrel:.2e}, W_shared_triu {wrel:.2e} (Frobenius, relative)")
    assert wrel <= 5e-5, (tag, wrel)


@pytest.mark.parametrize("rebalance", ["lazy", "exact"])
@pytest.mark.parametrize("name", sorted(HIER_PINS))
def test_hierarchical_model_matches_reference(golden, name, rebalance):
    """LocalHierarchicalKDE against the reference sweep by sweep
    (lakde/local_hierarchical_kde.py:198-311): initialisation; on the
    jitter-free hier_circle8_nojit every sweep's normalisers, nu, b and
    shared prior factor W_shared_triu at the 1e-4 / 1e-5 bars and the
    held-out mean within 1e-4; on the fixtures whose FP32 trajectory
    separates from the reference's (FP32 jitter decisions, rounding
    amplified by the shared-prior update), the pinned sweeps and the
    held-out mean within 2e-3, or 1.2x the reference's own FP32-vs-FP64
    spread where that is larger."""
    from lakde_amd import LocalHierarchicalKDE
    pinned, init_rtol, bar = HIER_PINS[name]
    d = golden(name)
    X, Y = _t(d["X"]), _t(d["Y"])
    t = float(d["threshold"])
    m = LocalHierarchicalKDE(k_or_knn_graph=_t(d["graph"]), block_size=int(d["block_size"]),
                             rebalance=rebalance)
    m.init_parameters_(X, t)
    m.init_summaries_(X, t)
    np.testing.assert_allclose(m.rnm_log_consts.cpu().numpy(), d["init_log_consts"],
                               rtol=init_rtol, atol=1e-4 if init_rtol <= 1e-4 else 0)
    np.testing.assert_allclose(m.nu.cpu().numpy(), d["init_nu"], rtol=1e-5)
    nb = (d["X"].shape[0] - 1) // m.block_size + 1
    for it in range(int(d["iterations"])):
        for b in range(nb):
            m.partial_step_(X, b, sparsity_threshold=t)
            m.iter_steps += 1
        if it < pinned:
            _check_hier_state_gpu(m, d, it, f"{name} sweep {it}")
    m.materialize()
    lp = m.log_pred_density(X, Y).cpu().numpy()
    ref = d["log_pred"]
    rel = abs(lp.mean() - ref.mean()) / abs(ref.mean())
    if bar is None:
        bar = max(2e-3, 1.2 * _fp64_spread(d, name))
    print(f"[{name}] mean log p {lp.mean():.6f} (reference {ref.mean():.6f}) rel {rel:.2e} "
          f"(bar {bar:.2e})")
    assert rel <= bar


@pytest.mark.parametrize("rebalance", ["lazy", "exact"])
@pytest.mark.parametrize("name,model", [("hier_latent_small", "hier"), ("hier_mix6_small", "hier"),
                                        ("knn_mix6_small", "knn")])
def test_fit_is_reproducible(golden, name, model, rebalance):
    """Two fits from the same inputs agree bit for bit: every per-column sum
    that feeds a normaliser is an order-independent fixed-point sum, so the
    order in which concurrent threads add to it does not matter (the
    hierarchical fixtures amplify any rounding difference through their
    jitter decisions and the shared-prior update)."""
    from lakde_amd import LocalHierarchicalKDE, LocalKNearestKDE
    cls = LocalHierarchicalKDE if model == "hier" else LocalKNearestKDE
    d = golden(name)
    outs = []
    for _ in range(2):
        m, lp = _fit(d, cls, rebalance)
        st = [lp, m.rnm_log_consts.cpu().numpy(), m.nu.cpu().numpy(),
              m.W_inv_jitter.cpu().numpy()]
        if model == "hier":
            st.append(m.W_shared_triu.cpu().numpy())
        rows, cols, vals = _pairs(m)
        st += [rows, cols, vals]
        outs.append(st)
    for k, (a, b) in enumerate(zip(*outs)):
        np.testing.assert_array_equal(a, b, err_msg=f"{name} {rebalance}: state item {k}")


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_prior_cache_matches_recomputed_prior(precision):
    """The per-fit kNN prior cache (prior + jitter and its factor, computed
    once) against recomputing the prior and its factorisation at every
    partial step and prediction, as lakde/local_knn_kde.py:195-207 does: the
    same factor enters the same kernels, so the fits agree bit for bit."""
    from lakde_amd import LocalKNearestKDE, synthetic
    X, Y = synthetic.make("mix64_d21", 3000, 300)
    graph = _t(O.compute_knn_graph(X, 32))
    Xd, Yd = _t(X), _t(Y)
    outs = []
    for cached in (True, False):
        m = LocalKNearestKDE(k_or_knn_graph=graph, block_size=512, precision=precision)
        m.prior_cache = cached
        m.fit(Xd, iterations=2, sparsity_threshold=1 / 32)
        assert (m._prior_cache is not None) == cached
        lp = m.log_pred_density(Xd, Yd)
        outs.append([lp.cpu().numpy(), m.rnm_log_consts.cpu().numpy(), m.nu.cpu().numpy(),
                     m.sigma_0_jitter.cpu().numpy(), m.logdet_sigma_0.cpu().numpy(),
                     m.sigma_0_lambda_mm_trace.cpu().numpy(), m.b.cpu().numpy()])
    for k, (a, b) in enumerate(zip(*outs)):
        np.testing.assert_array_equal(a, b, err_msg=f"state item {k}")


@pytest.mark.parametrize("name,model", [("state_knn_mix6", "knn"), ("state_hier_circle8", "hier")])
def test_reference_state_interop(golden, name, model):
    """Checkpoints between the reference and lakde_amd
    (lakde/kdes.py:115-135, local_adaptive_kde.py:272-309,
    local_knn_kde.py:343-367, local_hierarchical_kde.py:463-490).
    Forward: a state the reference wrote (its state_dict() after 2 sweeps,
    saved with torch.save and read here with torch.load(weights_only=True))
    loads into the lakde_amd model, and 2 more sweeps on the GPU reproduce
    the reference's continuation (held-out mean within 1e-4 relative,
    normalisers, nu, b within 1e-4, equal jitters, shared prior factor
    within 1e-5 of its largest entry).  Reverse: the lakde_amd state_dict,
    through torch.save and the same safe loader, has the reference's keys
    and loads into the oracle (the reference's algorithm on the CPU), whose
    next sweep matches the GPU model's next sweep at the same bars."""
    import io

    from lakde_amd import LocalHierarchicalKDE, LocalKNearestKDE
    from test_oracle_golden import _check_continuation, load_reference_state
    d = golden(name)
    sd = load_reference_state(name)
    X, Y = _t(d["X"]), _t(d["Y"])
    t = float(d["threshold"])
    cls, ocls = ((LocalKNearestKDE, O.LocalKNearestKDE) if model == "knn"
                 else (LocalHierarchicalKDE, O.LocalHierarchicalKDE))
    m = cls(block_size=int(d["block_size"]))
    m.load_state_dict(sd, X)
    m.fit(X, iterations=int(d["after"]), sparsity_threshold=t)
    lp = m.log_pred_density(X, Y).cpu().numpy()
    _check_continuation(lp, m.rnm_log_consts.cpu().numpy(), m.nu.cpu().numpy(),
                        m.b.cpu().numpy(), m.W_inv_jitter.cpu().numpy(), d, f"{name} forward",
                        m.W_shared_triu.cpu().numpy() if model == "hier" else None)
    # reverse
    buf = io.BytesIO()
    torch.save(m.state_dict(), buf)
    buf.seek(0)
    ours = torch.load(buf, weights_only=True, map_location="cpu")
    assert sorted(ours) == sorted(sd)
    mo = ocls(block_size=int(d["block_size"]))
    mo.load_state_dict(ours)
    mo.fit(d["X"], iterations=1, sparsity_threshold=t)
    m.fit(X, iterations=1, sparsity_threshold=t)
    ref = {"log_pred": mo.log_pred_density(d["X"], d["Y"]), "log_consts": mo.rnm_log_consts,
           "nu": mo.nu, "b": mo.b, "W_inv_jitter": mo.W_inv_jitter}
    if model == "hier":
        ref["W_shared_triu"] = mo.W_shared_triu
    _check_continuation(m.log_pred_density(X, Y).cpu().numpy(), m.rnm_log_consts.cpu().numpy(),
                        m.nu.cpu().numpy(), m.b.cpu().numpy(), m.W_inv_jitter.cpu().numpy(), ref,
                        f"{name} reverse", m.W_shared_triu.cpu().numpy() if model == "hier"
                        else None)


def test_knn_model_matches_oracle_trajectory():
    """Per-sweep log normalisers against the float32 oracle on a fresh
    seeded sample (not one of the reference fixtures)."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import synthetic
    X, Y = synthetic.make("mix32_d6", 3000, 400)
    graph = O.compute_knn_graph(X, 64)
    mo = O.LocalKNearestKDE(k_or_knn_graph=graph, block_size=1024)
    md = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=1024)
    Xd = _t(X)
    t = 1 / 64
    mo.init_parameters_(X, t)
    mo.init_summaries_(X, t)
    md.init_parameters_(Xd, t)
    md.init_summaries_(Xd, t)
    np.testing.assert_allclose(md.rnm_log_consts.cpu().numpy(), mo.rnm_log_consts, rtol=1e-5,
                               atol=1e-4)
    nb = (X.shape[0] - 1) // 1024 + 1
    for it in range(3):
        for b in range(nb):
            mo.partial_step_(X, b, sparsity_threshold=t)
            md.partial_step_(Xd, b, sparsity_threshold=t)
        np.testing.assert_allclose(md.rnm_log_consts.cpu().numpy(), mo.rnm_log_consts,
                                   rtol=1e-4, atol=1e-3)
        np.testing.assert_allclose(md.b.cpu().numpy(), mo.b, rtol=1e-4)
    lp_o = mo.log_pred_density(X, Y)
    lp_d = md.log_pred_density(Xd, _t(Y)).cpu().numpy()
    assert abs(lp_d.mean() - lp_o.mean()) <= 1e-4 * abs(lp_o.mean())


def test_state_dict_round_trip(golden):
    from lakde_amd import LocalKNearestKDE
    d = golden("knn_mix6_small")
    X, Y = _t(d["X"]), _t(d["Y"])
    m = LocalKNearestKDE(k_or_knn_graph=_t(d["graph"]), block_size=int(d["block_size"]))
    m.fit(X, iterations=2, sparsity_threshold=float(d["threshold"]))
    sd = m.state_dict()
    m2 = LocalKNearestKDE(block_size=int(d["block_size"]))
    m2.load_state_dict(sd, X)
    m.fit(X, iterations=1, sparsity_threshold=float(d["threshold"]))
    m2.fit(X, iterations=1, sparsity_threshold=float(d["threshold"]))
    a = m.log_pred_density(X, Y).cpu().numpy()
    b = m2.log_pred_density(X, Y).cpu().numpy()
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name,n,k", [("mix32_d6", 4000, 32), ("mix64_d21", 6000, 64)])
def test_lazy_rebalance_matches_exact_mid_sweep(name, n, k):
    """The lazy rebalance (per-column factors, folded once per sweep) against
    the exact whole-arena rebalance on the same inputs, read mid-sweep (the
    responsibilities handed out fold the pending factors) and after the fit:
    values within 1e-4 relative on the common entries, log normalisers within
    1e-5 relative, index sets equal up to a few threshold-crossing entries.
    mix64_d21 raises normalisers by more than the stored form's window
    (2^40) within single steps, so most of its lazy steps re-base columns
    (their other blocks' values made true relative to the step's normaliser,
    G = 1 / coeff), some of them after a lazy step of the same sweep
    (G != 1); the exact fallback, for columns that cannot be re-based, must
    not apply G to the step's own block, whose values the compaction wrote
    as true values.  There the two modes' float32 roundings are amplified
    (an entry moved across the threshold moves its kernel's posterior), so
    mix64_d21 is held to bounds on the differences instead: normalisers
    within 0.05 (1e-4 on average), values within 5e-3 (1e-4 on average), the
    held-out density within 1e-3 nats (1e-4 on average); with G applied to
    the step's block the normalisers differ by up to 0.64, 1.6e-2 on
    average."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import synthetic
    X, Y = synthetic.make(name, n, 300)
    graph = O.compute_knn_graph(X, k)
    t = 1 / k
    Xd = _t(X)
    snaps = {}
    flags = []

    def grab(tag):
        def cb(X_, model, step):
            if tag == "lazy":
                flags.append(int(model._flags[model.arena.num_blocks + 1]))
            if step == 5:
                snaps[tag] = (_pairs(model), model.rnm_log_consts.cpu().numpy())
        return cb

    ms = {}
    for reb in ("lazy", "exact"):
        m = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=512, rebalance=reb)
        m.fit(Xd, iterations=3, sparsity_threshold=t, callbacks=grab(reb))
        ms[reb] = m
    nbk = (n - 1) // 512 + 1
    exact = [(f & 1) != 0 for f in flags]
    rebased = [f == 2 for f in flags]
    # steps whose G can differ from 1: not the first of a sweep, and after
    # a lazy (not exact) step
    g_set = [i % nbk != 0 and not exact[i - 1] for i in range(len(flags))]
    print(f"[{name}] steps {len(flags)}: exact {sum(exact)} ({sum(e and g for e, g in zip(exact, g_set))} "
          f"with G != 1), re-based {sum(rebased)} ({sum(r and g for r, g in zip(rebased, g_set))} "
          f"with G != 1)")
    if name == "mix64_d21":
        assert sum(r and g for r, g in zip(rebased, g_set)) > 0
    strict = name == "mix32_d6"

    def bounded(a, b, mx, mean, what):
        d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
        print(f"[{name}] {what}: max {d.max():.2e}, mean {d.mean():.2e}")
        assert d.max() < mx and d.mean() < mean, (what, d.max(), d.mean())

    for stage in ("mid", "end"):
        if stage == "mid":
            (pl, cl), (pe, ce) = snaps["lazy"], snaps["exact"]
        else:
            pl, cl = _pairs(ms["lazy"]), ms["lazy"].rnm_log_consts.cpu().numpy()
            pe, ce = _pairs(ms["exact"]), ms["exact"].rnm_log_consts.cpu().numpy()
        if strict:
            _close_but_flips(cl, ce)
        else:
            bounded(cl, ce, 0.05, 1e-4, f"{stage} normalisers")
        N = X.shape[0]
        kl = pl[0] * N + pl[1]
        ke = pe[0] * N + pe[1]
        common, il, ie = np.intersect1d(kl, ke, return_indices=True)
        if strict:
            _close_but_flips(pl[2][il], pe[2][ie], rtol=1e-4, atol=1e-9, flip_tol=0.05)
        else:
            bounded(pl[2][il], pe[2][ie], 5e-3, 1e-4, f"{stage} values")
        # a float32 rounding difference can move an entry across the
        # threshold at its E-step; such entries are rare
        ndiff = kl.size + ke.size - 2 * common.size
        assert ndiff <= 1e-4 * common.size + 5, (stage, ndiff, common.size)
        print(f"[{stage}] common {common.size}, only lazy {kl.size - common.size}, "
              f"only exact {ke.size - common.size}")
    a = ms["lazy"].log_pred_density(Xd, _t(Y)).cpu().numpy()
    b = ms["exact"].log_pred_density(Xd, _t(Y)).cpu().numpy()
    if strict:
        _close_but_flips(a, b, flip_tol=1e-3)
    else:
        bounded(a, b, 1e-3, 1e-4, "held-out density")


def test_estep_skip_matches_full_evaluation():
    """The pair-skipping E-step (Morton-ordered column chunks, kernels bounded
    on each chunk's box and skipped when provably tens of nats below the
    normaliser) against the full evaluation on the same inputs: log
    normalisers within 1e-5 relative, values within 1e-4 relative on the
    common entries, index sets equal up to a few threshold-crossing entries,
    identical posterior jitters and the same held-out density."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import synthetic
    X, Y = synthetic.make("mix32_d6", 20000, 500)
    graph = O.compute_knn_graph(X, 64)
    t = 1 / 64
    Xd = _t(X)
    ms = {}
    for skip in (True, False):
        m = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=512, estep_skip=skip)
        m.fit(Xd, iterations=3, sparsity_threshold=t, callbacks=[])
        ms[skip] = m
    assert ms[True]._col_order is not None and ms[False]._col_order is None
    cs = ms[True].rnm_log_consts.cpu().numpy()
    cf = ms[False].rnm_log_consts.cpu().numpy()
    _close_but_flips(cs, cf)
    ps, pf = _pairs(ms[True]), _pairs(ms[False])
    N = X.shape[0]
    ks = ps[0] * N + ps[1]
    kf = pf[0] * N + pf[1]
    common, i_s, i_f = np.intersect1d(ks, kf, return_indices=True)
    _close_but_flips(ps[2][i_s], pf[2][i_f], rtol=1e-4, atol=1e-9, flip_tol=0.05)
    ndiff = ks.size + kf.size - 2 * common.size
    assert ndiff <= 1e-4 * common.size + 5, (ndiff, common.size)
    js, jf = ms[True].W_inv_jitter.cpu().numpy(), ms[False].W_inv_jitter.cpu().numpy()
    assert (js != jf).sum() <= 2
    _close_but_flips(ms[True].nu.cpu().numpy(), ms[False].nu.cpu().numpy(), rtol=1e-5, atol=0,
                     flip_tol=0.5)
    a = ms[True].log_pred_density(Xd, _t(Y)).cpu().numpy()
    b = ms[False].log_pred_density(Xd, _t(Y)).cpu().numpy()
    _close_but_flips(a, b, flip_tol=1e-3)
    print(f"common {common.size}, differing {ndiff}")


@pytest.mark.parametrize("D,threshold", [(12, 1 / 32), (3, 0.0)])
def test_estep_skip_matches_full_evaluation_other_shapes(D, threshold):
    """Pair skipping against full evaluation at another dimension (the
    largest skipping kernels are instantiated up to D=16) and with
    threshold 0 (keep r != 0: the skip margin is 110 nats), on a ragged last
    block (N not a multiple of the block size)."""
    from lakde_amd import LocalKNearestKDE
    rng = np.random.default_rng(D)
    centres = rng.standard_normal((16, D)) * 4
    X = (centres[rng.integers(0, 16, 6000)] + rng.standard_normal((6000, D))).astype(np.float32)
    Y = (centres[rng.integers(0, 16, 300)] + rng.standard_normal((300, D))).astype(np.float32)
    graph = O.compute_knn_graph(X, 24)
    Xd = _t(X)
    ms = {}
    for skip in (True, False):
        m = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=512, estep_skip=skip)
        m.fit(Xd, iterations=2, sparsity_threshold=threshold, callbacks=[])
        ms[skip] = m
    _close_but_flips(ms[True].rnm_log_consts.cpu().numpy(),
                     ms[False].rnm_log_consts.cpu().numpy())
    ps, pf = _pairs(ms[True]), _pairs(ms[False])
    N = X.shape[0]
    ks, kf = ps[0] * N + ps[1], pf[0] * N + pf[1]
    common, i_s, i_f = np.intersect1d(ks, kf, return_indices=True)
    _close_but_flips(ps[2][i_s], pf[2][i_f], rtol=1e-4, atol=1e-9, flip_tol=0.05)
    assert ks.size + kf.size - 2 * common.size <= 1e-4 * common.size + 5
    a = ms[True].log_pred_density(Xd, _t(Y)).cpu().numpy()
    b = ms[False].log_pred_density(Xd, _t(Y)).cpu().numpy()
    _close_but_flips(a, b, flip_tol=1e-3)


def test_column_order_chunks_and_boxes():
    """utils.column_order: a permutation whose 64-column chunks have the
    returned boxes as their exact coordinate-wise min / max."""
    from lakde_amd.utils import column_order
    rng = np.random.default_rng(2)
    X = rng.standard_normal((1000, 5)).astype(np.float32)
    perm, box = column_order(_t(X))
    perm, box = perm.cpu().numpy(), box.cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(1000))
    assert box.shape == (16, 2, 5)
    for c in range(16):
        pts = X[perm[c * 64:(c + 1) * 64]]
        np.testing.assert_array_equal(box[c, 0], pts.min(0))
        np.testing.assert_array_equal(box[c, 1], pts.max(0))


def test_arena_growth_inside_sweep_matches():
    """A step whose entries outgrow the arena inside the native sweep loop
    (the driver returns, the arena grows, the step is finished by the
    resume entry point) gives the same fit as an arena that never grows:
    normalisers within 1e-5 relative (float atomics in the column sums make
    the last bit order-dependent), entry counts within 1e-3."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import _fused as F
    from lakde_amd import synthetic
    X, Y = synthetic.make("mix32_d6", 3000, 100)
    graph = _t(O.compute_knn_graph(X, 32))
    Xd = _t(X)
    t = 1 / 32
    ref = LocalKNearestKDE(k_or_knn_graph=graph, block_size=256)
    ref.fit(Xd, iterations=2, sparsity_threshold=t)
    m = LocalKNearestKDE(k_or_knn_graph=graph, block_size=256)
    m.init_parameters_(Xd, t)
    m.init_summaries_(Xd, t)
    # shrink the arena to the current largest block: later steps overflow
    old = m.arena
    cap = max(old.nnz)
    new = F.Arena(old.num_blocks, old.block_size, cap, Xd.device)
    for name in ("row", "col", "val"):
        getattr(new, name).copy_(getattr(old, name)[:, :cap])
    new.rowptr.copy_(old.rowptr)
    new.nnz_dev.copy_(old.nnz_dev)
    new.nnz = list(old.nnz)
    m.arena = new
    m._step_cache = {}
    m.iter_steps = 1  # initialised: fit() continues instead of initialising again
    m.fit(Xd, iterations=2, sparsity_threshold=t)
    print(f"cap {cap} -> {m.arena.cap}")
    assert m.arena.cap > cap
    np.testing.assert_allclose(m.rnm_log_consts.cpu().numpy(), ref.rnm_log_consts.cpu().numpy(),
                               rtol=1e-5, atol=1e-5)
    assert abs(m.arena.total_nnz() - ref.arena.total_nnz()) <= 1e-3 * ref.arena.total_nnz()
