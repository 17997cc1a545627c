"""Precision form against the FP32 GPU path and the FP64 oracle.

The fit's E-step is the screening form (every pair screened on the BF16
matrix cores with a rigorous error bound, the pairs that can matter
recomputed exactly in FP32); the bf16x3 E-step and predictive kernels
(split-precision products on the matrix cores) are tested beside it.

Bars, written per assertion: screening normalisers within float32
summation-order differences of the FP32 kernels (1e-6 relative) and its
FIXED-mode CSR identical; the bf16x3 E-step normalisers within 2e-2 nats of
the FP32 kernels per column; the fitted model's held-out log density within
1e-3 nats/point (mean absolute) of the FP32 path and of the FP64 oracle;
the bf16x3 predictive kernel within 1e-2 nats per point of the FP32 kernel
on the same posterior; the screening predictive density within 1e-4 nats
per point of the FP32 kernel and of float64 arithmetic on the same records.
"""

import numpy as np
import pytest
import torch

from oracle import lakde_np as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _fit(X, Y, graph, precision, bs, t, iters):
    from lakde_amd import LocalKNearestKDE
    m = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=bs, precision=precision)
    Xd = _t(X)
    m.fit(Xd, iterations=iters, sparsity_threshold=t)
    return m, m.log_pred_density(Xd, _t(Y)).cpu().numpy()


@pytest.mark.parametrize("name,n,m,k,tol", [("mix64_d21", 3000, 300, 32, 2e-2),
                                            ("mix128_d63", 3000, 200, 96, 2e-2),
                                            ("latent_d43", 2000, 200, 32, None)])
def test_bf16_estep_lse_matches_fp32(name, n, m, k, tol):
    """latent_d43 has integer columns: its posteriors are near singular
    (factor entries ~1e3), where bf16x3's 2^-17 relative product error is
    amplified; it is reported with a looser bound."""
    from lakde_amd import _fused as F
    from lakde_amd._lib import ESTEP_ACCUM
    from lakde_amd import synthetic
    X, Y = synthetic.make(name, n, m)
    graph = O.compute_knn_graph(X, k)
    model, _ = _fit(X, Y, graph, "fp32", 512, 1 / k, 1)
    Xd = _t(X)
    N, D = X.shape
    s, e = 0, 512
    rec = model.rec[s:e]
    c32 = torch.full((N,), float("-inf"), device=DEV)
    cbf = torch.full((N,), float("-inf"), device=DEV)
    ws = F.EstepWorkspace(N, e - s, Xd.device)
    F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, c32, ws)
    frag = torch.empty((e - s) * F.bf16_frag_bytes(D), dtype=torch.uint8, device=DEV)
    F.pack_frag_bf16(rec, D, frag)
    F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, cbf, ws, bf16=frag)
    d = (cbf - c32).abs().cpu().numpy()
    print(f"[{name}] block-0 LSE |bf16 - fp32|: max {d.max():.3e} mean {d.mean():.3e} nats")
    assert np.all(np.isfinite(cbf.cpu().numpy()))
    if tol is not None:
        assert d.max() < tol
    else:
        assert d.mean() < 0.5


def _screen_case(name, n, k):
    from lakde_amd import synthetic
    if name == "normal8":
        rng = np.random.default_rng(8)
        X = rng.standard_normal((n, 8)).astype(np.float32)
        Y = rng.standard_normal((100, 8)).astype(np.float32)
        return X, Y
    return synthetic.make(name, n, 100)


@pytest.mark.parametrize("name,n,k,t", [("mix64_d21", 3000, 32, 1 / 32),
                                        ("mix128_d63", 3000, 96, 1 / 96),
                                        ("latent_d43", 2000, 32, 1 / 32),
                                        ("normal8", 2500, 64, 0.0)])
def test_screen_estep_matches_fp32_kernels(name, n, k, t):
    """The screening E-step (BF16 matrix-core screen with a rigorous error
    bound, exact FP32 recompute of the surviving pairs) against the FP32
    E-step kernel on the same records: the ACCUM normalisers within float32
    summation-order differences (1e-6 relative), and the FIXED-mode CSR of
    the block identical entry for entry (the keep decisions of both run on
    the same exact FP32 values).  normal8 with threshold 0 keeps every
    nonzero responsibility (110-nat margin): its columns queue more pairs
    than a lane holds and more keep candidates than the keep queue, so the
    resume passes and the keep pass run."""
    from lakde_amd import _fused as F
    from lakde_amd._lib import ESTEP_ACCUM, ESTEP_FIXED
    X, Y = _screen_case(name, n, k)
    graph = O.compute_knn_graph(X, k)
    model, _ = _fit(X, Y, graph, "fp32", 512, t if t else 1 / k, 1)
    Xd = _t(X)
    N, D = X.shape
    pts = F.ScreenPoints(Xd)
    for b in (0, (N - 1) // 512):
        s, e = b * 512, min(N, b * 512 + 512)
        nb = e - s
        print(f"[{name}] block {b}: rows {s}..{e} D={D} threshold={t}", flush=True)
        rec = model.rec[s:e]
        frag = torch.empty(nb * F.screen_frag_bytes(D), dtype=torch.uint8, device=DEV)
        F.pack_screen(rec, D, pts.origin, frag)
        ws = F.EstepWorkspace(N, nb, Xd.device)
        c32 = torch.full((N,), float("-inf"), device=DEV)
        csc = c32.clone()
        # start from the model's normalisers minus a little, so that the
        # ACCUM update is a real combination of both terms
        base = model.rnm_log_consts - 3.0
        c32.copy_(base)
        csc.copy_(base)
        F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, c32, ws)
        F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, csc, ws, screen=(frag, pts))
        np.testing.assert_allclose(csc.cpu().numpy(), c32.cpu().numpy(), rtol=1e-6, atol=1e-5)
        # FIXED: same normaliser in, identical CSR out
        outs = []
        for use_screen in (False, True):
            arena = F.Arena(1, nb, 1 << 21, Xd.device)
            F.estep(Xd, rec, s, ESTEP_FIXED, t, c32, ws, arena=arena, b=0,
                    screen=(frag, pts) if use_screen else None)
            cnt = int(arena.nnz_dev[0])
            assert cnt <= arena.cap
            outs.append((arena.rowptr[0, :nb + 1].cpu().numpy(), arena.col[0, :cnt].cpu().numpy(),
                         arena.val[0, :cnt].cpu().numpy(), arena.row[0, :cnt].cpu().numpy()))
        (p0, c0, v0, r0), (p1, c1, v1, r1) = outs
        print(f"[{name}] block {b}: {c0.size} kept entries", flush=True)
        np.testing.assert_array_equal(p1, p0)
        np.testing.assert_array_equal(r1, r0)
        np.testing.assert_array_equal(c1, c0)
        np.testing.assert_array_equal(v1, v0)
        # every stored responsibility passes the reference's threshold test
        # (lakde/local_adaptive_kde.py:139-143: threshold_ zeroes r <= t, nonzero
        # keeps the rest; t = 0 keeps r > 0)
        assert np.all(v1 > t), (name, b, float(v1.min()))


@pytest.mark.parametrize("name,n,k", [("mix128_d63", 3000, 96), ("latent_d43", 2500, 32)])
def test_screen_init_matches_fp32(name, n, k):
    """The initialisation of the precision form (log-sum-exp passes started
    from the graph-neighbour bound of every normaliser, then the keep pass)
    against the FP32 path's: the normalisers within float32 summation-order
    differences, the initial CSR identical entry for entry, and the bound
    below every final normaliser."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import _fused as F
    X, _ = _screen_case(name, n, k)
    graph = O.compute_knn_graph(X, k)
    Xd = _t(X)
    t = 1 / k
    ms = {}
    for prec in ("fp32", "bf16"):
        m = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=512, precision=prec)
        m.init_parameters_(Xd, t)
        m.init_summaries_(Xd, t)
        ms[prec] = m
    a, b = ms["fp32"], ms["bf16"]
    np.testing.assert_allclose(b.rnm_log_consts.cpu().numpy(), a.rnm_log_consts.cpu().numpy(),
                               rtol=1e-6, atol=1e-5)
    # the normalisers differ in the last bits, so a value at the threshold
    # may be kept by one form only (values after the column renormalisation)
    flips, total = 0, 0
    a.materialize()
    b.materialize()
    for blk in range(a.arena.num_blocks):
        ent = []
        for mm in (a, b):
            nz = mm.arena.nnz[blk]
            r = mm.arena.row[blk, :nz].cpu().numpy().astype(np.int64)
            c = mm.arena.col[blk, :nz].cpu().numpy().astype(np.int64)
            v = mm.arena.val[blk, :nz].cpu().numpy()
            ent.append(dict(zip((r << 32 | c).tolist(), v.tolist())))
        ea, eb = ent
        for key in set(ea) ^ set(eb):
            v = ea.get(key, eb.get(key))
            assert v <= 2 * t, f"block {blk}: entry {key} value {v} kept by one form"
            flips += 1
        total += len(ea)
        common = sorted(set(ea) & set(eb))
        np.testing.assert_allclose([eb[q] for q in common], [ea[q] for q in common],
                                   rtol=1e-5, atol=1e-9)
    print(f"[{name}] entries kept by one form only: {flips} of {total}")
    assert flips <= max(2, 1e-4 * total)
    pts = F.ScreenPoints(Xd)
    lb = pts.knn_bound(b.rec, Xd, b.knn_graph).cpu().numpy()
    c = b.rnm_log_consts.cpu().numpy()
    print(f"[{name}] normaliser - bound: min {np.min(c - lb):.3e} median "
          f"{np.median(c - lb):.3f} nats")
    assert np.all(lb <= c + 1e-5 * np.abs(c))


@pytest.mark.parametrize("name,n,m,k,iters", [("mix64_d21", 4000, 400, 32, 2),
                                              ("mix128_d63", 2000, 200, 96, 2)])
def test_bf16_fit_matches_fp32_and_fp64(name, n, m, k, iters):
    """k >= D so the kNN covariances are full rank (with k < D they are
    singular, the jitter is eps-sized and float32 and float64 fits differ by
    hundreds of nats, whatever the quadratic-form precision)."""
    from lakde_amd import synthetic
    X, Y = synthetic.make(name, n, m)
    graph = O.compute_knn_graph(X, k)
    t = 1 / k
    _, lp32 = _fit(X, Y, graph, "fp32", 1024, t, iters)
    _, lpbf = _fit(X, Y, graph, "bf16", 1024, t, iters)
    mo = O.LocalKNearestKDE(k_or_knn_graph=graph, block_size=1024, dtype=np.float64)
    mo.fit(X, iterations=iters, sparsity_threshold=t)
    lp64 = mo.log_pred_density(X, Y)
    e_bf_32 = np.abs(lpbf - lp32)
    e_bf_64 = np.abs(lpbf - lp64)
    e_32_64 = np.abs(lp32 - lp64)
    print(f"[{name}] |bf16-fp32| mean {e_bf_32.mean():.2e} max {e_bf_32.max():.2e}; "
          f"|bf16-fp64| mean {e_bf_64.mean():.2e} max {e_bf_64.max():.2e}; "
          f"|fp32-fp64| mean {e_32_64.mean():.2e} max {e_32_64.max():.2e} nats/pt")
    assert e_bf_32.mean() < 1e-3
    assert e_bf_64.mean() < 1e-3


def test_bf16_predictive_matches_fp32():
    from lakde_amd import _fused as F
    from lakde_amd import synthetic
    X, Y = synthetic.make("mix128_d63", 3000, 500)
    graph = O.compute_knn_graph(X, 96)
    model, lp32 = _fit(X, Y, graph, "fp32", 1024, 1 / 96, 1)
    N, D = X.shape
    frag = torch.empty(N * F.bf16_frag_bytes(D), dtype=torch.uint8, device=DEV)
    F.pack_frag_bf16(model.rec, D, frag)
    Yd = _t(Y)
    out32 = F.pred_density(model.rec, Yd, torch.empty(Y.shape[0], device=DEV)).cpu().numpy()
    outbf = F.pred_density_bf16(model.rec, frag, Yd,
                                torch.empty(Y.shape[0], device=DEV)).cpu().numpy()
    d = np.abs(outbf - out32)
    print(f"predictive |bf16 - fp32|: max {d.max():.3e} mean {d.mean():.3e} nats")
    assert d.max() < 1e-2
    # the model's precision-form predictive density, bf16x3 (default) and
    # screened (pred_screen), on the same fitted state: both against the
    # FP32 kernel's values (log_pred_density recomputes the same posteriors)
    model.precision = "bf16"
    lp_x3 = model.log_pred_density(_t(X), Yd).cpu().numpy()
    model.pred_screen = True
    lp_sc = model.log_pred_density(_t(X), Yd).cpu().numpy()
    assert np.abs(lp_x3 - out32).max() < 1e-2
    assert np.abs(lp_sc - out32).max() < 1e-4


def _pred64(rec, D, Y):
    """log p(y) in float64 arithmetic from the same float32 records
    (lakde/local_adaptive_kde.py:209-223)."""
    r = rec.double().cpu().numpy()
    N = r.shape[0]
    P = D * (D + 1) // 2
    ti = np.tril_indices(D)
    L = np.zeros((N, D, D))
    L[:, ti[0], ti[1]] = r[:, :P]
    xc, coeff, h = r[:, P:P + D], r[:, P + D + 2], r[:, P + D + 3]
    Y64 = np.asarray(Y, np.float64)
    out = np.empty(Y64.shape[0])
    for i in range(Y64.shape[0]):
        z = np.einsum("nij,nj->ni", L, Y64[i][None] - xc)
        t = coeff - h * np.log1p((z * z).sum(1))
        mx = t.max()
        out[i] = mx + np.log(np.exp(t - mx).sum()) - np.log(N)
    return out


@pytest.mark.parametrize("name,n,m,k", [("mix64_d21", 3000, 300, 32),
                                        ("mix128_d63", 6000, 150, 96),
                                        ("latent_d43", 2500, 77, 32),
                                        ("mix64_d21", 21000, 1, 32)])
def test_screen_predictive_matches_fp32(name, n, m, k):
    """The screening predictive density (BF16 screen with a rigorous bound,
    exact FP32 terms for the pairs within the margin of the running bound)
    against the FP32 predictive kernel on the same records: within 1e-4
    nats per point (float32 summation order; the dropped terms are below
    1e-6 of each split's partial sum), and both within 1e-4 relative
    of the same density in float64 arithmetic.  The shapes cover one and many
    kernel splits (21,000 kernels, one test point: 42 splits of 500) and
    partial workgroups of test points."""
    from lakde_amd import _fused as F
    from lakde_amd import synthetic
    X, Y = synthetic.make(name, n, m)
    graph = O.compute_knn_graph(X, k)
    model, _ = _fit(X, Y, graph, "fp32", 512, 1 / k, 1)
    N, D = X.shape
    Xd, Yd = _t(X), _t(Y)
    pts_x = F.ScreenPoints(Xd)
    frag = torch.empty(N * F.screen_frag_bytes(D), dtype=torch.uint8, device=DEV)
    F.pack_screen(model.rec, D, pts_x.origin, frag)
    pts_y = F.ScreenPoints(Yd, origin=pts_x.origin)
    out32 = F.pred_density(model.rec, Yd, torch.empty(m, device=DEV)).cpu().numpy()
    outsc = F.pred_density_screen(model.rec, frag, Yd, pts_y,
                                  torch.empty(m, device=DEV)).cpu().numpy()
    ref = _pred64(model.rec, D, Y[:min(m, 60)])
    d = np.abs(outsc - out32)
    print(f"[{name}] N={N} M={m}: |screen - fp32| max {d.max():.2e}, "
          f"|screen - float64| max {np.abs(outsc[:ref.size] - ref).max():.2e}", flush=True)
    assert np.all(np.isfinite(outsc))
    assert d.max() <= 1e-4
    # float32 arithmetic against float64: 1e-4 relative (latent_d43's
    # near-singular posteriors put the FP32 kernel itself 1.5e-4 nats away)
    bar = 1e-4 * np.maximum(1.0, np.abs(ref))
    assert np.all(np.abs(outsc[:ref.size] - ref) <= bar)
    assert np.all(np.abs(out32[:ref.size] - ref) <= bar)


def _csr_of(arena, nb):
    cnt = int(arena.nnz_dev[0])
    assert cnt <= arena.cap
    return (arena.rowptr[0, :nb + 1].cpu().numpy(), arena.row[0, :cnt].cpu().numpy(),
            arena.col[0, :cnt].cpu().numpy(), arena.val[0, :cnt].cpu().numpy())


@pytest.mark.parametrize("name,n,k,t", [("mix128_d63", 3000, 96, 1 / 96),
                                        ("mix64_d21", 3000, 32, 1 / 32),
                                        ("normal8", 2500, 64, 0.0)])
def test_screen_split_update_matches_unsplit(name, n, k, t):
    """The split form of the screening UPDATE E-step (the block's kernels in
    ranges of ~128 screened by separate workgroups, then a per-column
    combine and a keep pass for overflowing columns) against the unsplit
    kernel from the same fitted state: normalisers within float32
    summation-order differences (1e-6 relative), the kept entries identical
    up to values within 1e-5 of the threshold (the normalisers differ in
    their last bits) and their values within the normalisers' difference
    (r = exp(x - c)) plus 4 (|c| + |x - c|) 2^-23 relative.  normal8
    with threshold 0 overflows the keep queues, so the keep pass runs."""
    from lakde_amd import _fused as F
    from lakde_amd._lib import ESTEP_UPDATE
    X, Y = _screen_case(name, n, k)
    graph = O.compute_knn_graph(X, k)
    model, _ = _fit(X, Y, graph, "fp32", 512, t if t else 1 / k, 1)
    model.materialize()
    Xd = _t(X)
    N, D = X.shape
    pts = F.ScreenPoints(Xd)
    split = F.screen_split_workspace(N, 512, Xd.device)
    for b in (0, (N - 1) // 512):
        s, e = b * 512, min(N, b * 512 + 512)
        nb = e - s
        rec = model.rec[s:e]
        frag = torch.empty(nb * F.screen_frag_bytes(D), dtype=torch.uint8, device=DEV)
        F.pack_screen(rec, D, pts.origin, frag)
        ws = F.EstepWorkspace(N, nb, Xd.device)
        ind, val = model.arena.block(b)
        old = torch.zeros(N, device=DEV)
        old.index_add_(0, ind[1].long(), val)
        res = []
        for sp in (None, split):
            c = torch.empty(N, device=DEV)
            arena = F.Arena(1, nb, 1 << 22, Xd.device)
            F.estep(Xd, rec, s, ESTEP_UPDATE, t, c, ws, c_old=model.rnm_log_consts,
                    old_colsum=old, arena=arena, b=0, screen=(frag, pts), split=sp)
            res.append((c.cpu().numpy(), _csr_of(arena, nb)))
        np.testing.assert_allclose(res[1][0], res[0][0], rtol=1e-6, atol=1e-5)
        (pa, ra, ca, va), (pb, rb, cb, vb) = res[0][1], res[1][1]
        cmax = float(np.abs(res[0][0]).max())
        ea = dict(zip((ra.astype(np.int64) << 32 | ca).tolist(), va.tolist()))
        eb = dict(zip((rb.astype(np.int64) << 32 | cb).tolist(), vb.tolist()))
        flips = set(ea) ^ set(eb)
        for key in flips:
            v = ea.get(key, eb.get(key))
            assert abs(v - t) <= 1e-5 * max(t, 1e-30), (name, b, key, v)
        common = sorted(set(ea) & set(eb))
        a = np.array([ea[q] for q in common])
        bb = np.array([eb[q] for q in common])
        # r = exp(x - c): the values move by the normalisers' own difference,
        # and d = x - c is rounded to float32 (an ulp of |d| <= -log r_min)
        dc = float(np.abs(res[1][0] - res[0][0]).max())
        dmax = float(-np.log(max(a.min(), 1e-45)))
        np.testing.assert_allclose(bb, a, rtol=1.5 * dc + 4 * (cmax + dmax) * 2.0 ** -23 + 1e-7,
                                   atol=0)
        print(f"[{name}] block {b}: nb={nb}, {len(ea)} kept entries, {len(flips)} at the "
              f"threshold", flush=True)
