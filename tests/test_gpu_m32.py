"""The plain FP32 E-step on the matrix cores (v_mfma_f32_32x32x2_f32,
17 <= D <= 63) against the FP32 vector E-step kernel on the same records.

Bars: normalisers (ACCUM, UPDATE) within float32 rounding-order differences
(1e-6 relative, 1e-5 absolute floor: both sum every pair's FP32 log rho in
the same kernel order, with products and sums rounded differently); the
FIXED-mode CSR identical entry for entry (same normaliser in: the keep
decisions of both run on the exact values the compaction stores); the
UPDATE-mode CSR identical up to entries within 1e-5 relative of the
threshold, values within 4 |c| 2^-23 relative (the normalisers differ in
the last bits); the fitted model's held-out log density within 1e-5
relative of the vector path's.
"""

import numpy as np
import pytest
import torch

from oracle import lakde_np as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _case(name, n, k):
    from lakde_amd import synthetic
    if name == "normal17":
        rng = np.random.default_rng(17)
        return rng.standard_normal((n, 17)).astype(np.float32), None
    return synthetic.make(name, n, 100)


def _csr(arena, nb):
    cnt = int(arena.nnz_dev[0])
    assert cnt <= arena.cap
    return (arena.rowptr[0, :nb + 1].cpu().numpy(), arena.row[0, :cnt].cpu().numpy(),
            arena.col[0, :cnt].cpu().numpy(), arena.val[0, :cnt].cpu().numpy())


def _same_csr(a, b, t, tag, rtol):
    """rtol: of the values (0 with the same normaliser; an UPDATE's
    normalisers differ in the last bits, which moves r = exp(x - c) by
    |c| 2^-23 relative)."""
    (pa, ra, ca, va), (pb, rb, cb, vb) = a, b
    if np.array_equal(pa, pb) and np.array_equal(ca, cb):
        np.testing.assert_array_equal(ra, rb, err_msg=tag)
        np.testing.assert_allclose(vb, va, rtol=rtol, atol=0, err_msg=tag)
        return 0
    ea = dict(zip((ra.astype(np.int64) << 32 | ca).tolist(), va.tolist()))
    eb = dict(zip((rb.astype(np.int64) << 32 | cb).tolist(), vb.tolist()))
    for key in set(ea) ^ set(eb):
        v = ea.get(key, eb.get(key))
        assert abs(v - t) <= 1e-5 * t, f"{tag}: entry {key} value {v} kept by one kernel only"
    return len(set(ea) ^ set(eb))


@pytest.mark.parametrize("name,n,k,t", [("mix64_d21", 3000, 32, 1 / 32),
                                        ("mix128_d63", 3000, 96, 1 / 96),
                                        ("latent_d43", 2000, 32, 1 / 32),
                                        ("normal17", 2500, 64, 0.0)])
def test_m32_estep_matches_vector_kernel(name, n, k, t):
    """normal17 with threshold 0 keeps every nonzero responsibility: each
    column queues more keep candidates than a lane holds, so the UPDATE
    launch takes its keep pass and the FIXED launch flushes its queues."""
    from lakde_amd import LocalKNearestKDE
    from lakde_amd import _fused as F
    from lakde_amd._lib import ESTEP_ACCUM, ESTEP_FIXED, ESTEP_UPDATE
    X, _ = _case(name, n, k)
    graph = O.compute_knn_graph(X, k)
    model = LocalKNearestKDE(k_or_knn_graph=_t(graph), block_size=512)
    model.fp32_mfma = False
    Xd = _t(X)
    model.fit(Xd, iterations=1, sparsity_threshold=t if t else None)
    model.materialize()
    N, D = X.shape
    pts = F.ScreenPoints(Xd)
    for b in (0, (N - 1) // 512):
        s, e = b * 512, min(N, b * 512 + 512)
        nb = e - s
        print(f"[{name}] block {b}: rows {s}..{e} D={D} threshold={t}", flush=True)
        rec = model.rec[s:e]
        frag = torch.empty(nb * F.m32_frag_bytes(D), dtype=torch.uint8, device=DEV)
        F.pack_m32(rec, D, pts.origin, frag)
        ws = F.EstepWorkspace(N, nb, Xd.device)
        base = model.rnm_log_consts - 3.0
        cv, cm = base.clone(), base.clone()
        F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, cv, ws)
        F.estep(Xd, rec, s, ESTEP_ACCUM, 0.0, cm, ws, m32=(frag, pts))
        np.testing.assert_allclose(cm.cpu().numpy(), cv.cpu().numpy(), rtol=1e-6, atol=1e-5)
        # FIXED: same normaliser in
        outs = []
        for use in (False, True):
            arena = F.Arena(1, nb, 1 << 21, Xd.device)
            F.estep(Xd, rec, s, ESTEP_FIXED, t, cv, ws, arena=arena, b=0,
                    m32=(frag, pts) if use else None)
            outs.append(_csr(arena, nb))
        _same_csr(outs[0], outs[1], t, f"{name} block {b} FIXED", 0)
        # UPDATE from the fitted state: the block's old column sums
        ind, val = model.arena.block(b)
        old = torch.zeros(N, device=DEV)
        old.index_add_(0, ind[1].long(), val)
        res = []
        for use in (False, True):
            c = torch.empty(N, device=DEV)
            arena = F.Arena(1, nb, 1 << 21, Xd.device)
            F.estep(Xd, rec, s, ESTEP_UPDATE, t, c, ws, c_old=model.rnm_log_consts,
                    old_colsum=old, arena=arena, b=0, m32=(frag, pts) if use else None)
            res.append((c.cpu().numpy(), _csr(arena, nb)))
        np.testing.assert_allclose(res[1][0], res[0][0], rtol=1e-6, atol=1e-5)
        cmax = float(np.abs(res[0][0]).max())
        flips = _same_csr(res[0][1], res[1][1], t, f"{name} block {b} UPDATE",
                          4 * cmax * 2.0 ** -23)
        print(f"[{name}] block {b}: {res[0][1][1].size} kept entries, {flips} at the threshold",
              flush=True)


@pytest.mark.parametrize("name,n,k", [("mix64_d21", 4000, 32), ("mix128_d63", 2500, 96)])
def test_m32_fit_matches_vector_fit(name, n, k):
    from lakde_amd import LocalKNearestKDE, synthetic
    X, Y = synthetic.make(name, n, 300)
    graph = _t(O.compute_knn_graph(X, k))
    Xd, Yd = _t(X), _t(Y)
    lps = []
    for mfma in (False, True):
        m = LocalKNearestKDE(k_or_knn_graph=graph, block_size=512)
        m.fp32_mfma = mfma
        m.fit(Xd, iterations=2, sparsity_threshold=1 / k)
        assert m._uses_m32() == mfma
        lps.append(m.log_pred_density(Xd, Yd).cpu().numpy())
    d = np.abs(lps[1] - lps[0])
    print(f"[{name}] |m32 - vector| held-out: mean {d.mean():.2e} max {d.max():.2e} nats")
    assert abs(lps[1].mean() - lps[0].mean()) <= 1e-5 * abs(lps[0].mean())
    np.testing.assert_allclose(lps[1], lps[0], rtol=1e-4, atol=1e-3)
