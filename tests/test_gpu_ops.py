"""GPU parity of the reference ops and the fused kernels against the CPU
oracle (oracle/lakde_np.py) and the reference's golden vectors.

Bars: bit-exact for index work (ind2ptr, CSR indices); float32 results within
the tolerances written next to each assertion, measured against float64.
"""

import math

import numpy as np
import pytest
import torch

from oracle import lakde_np as O

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def K():
    from lakde_amd import kernels
    return kernels


# ------------------------------------------------------------ ind2ptr
@pytest.mark.parametrize("in_dt,out_dt", [(torch.int32, torch.int32), (torch.int32, torch.int64),
                                          (torch.int64, torch.int32), (torch.int64, torch.int64)])
def test_ind2ptr_bit_exact(K, in_dt, out_dt):
    rng = np.random.default_rng(1)
    for M, nnz in [(1, 0), (5, 0), (7, 1), (513, 40_000), (3000, 17)]:
        row = np.sort(rng.integers(0, M, size=nnz))
        ind = _t(row, in_dt)
        out = torch.empty(M + 1, dtype=out_dt, device=DEV)
        K.ind2ptr(ind, M, out=out)
        np.testing.assert_array_equal(_np(out), O.ind2ptr(row, M))
    r = _t(np.array([2, 2, 2, 5, 5, 6, 9, 9, 9, 9]), in_dt)
    out = K.ind2ptr(r, 12)
    assert out.dtype == in_dt
    np.testing.assert_array_equal(_np(out), [0, 0, 0, 3, 3, 3, 5, 6, 6, 6, 10, 10, 10])


def test_ind2ptr_golden(K, golden):
    g = golden("ops")
    np.testing.assert_array_equal(_np(K.ind2ptr(_t(g["i2p_row"]), 12)), g["i2p_ptr"])


# ------------------------------------------------------------ W_inv
@pytest.mark.parametrize("fp", [torch.float32, torch.float64])
@pytest.mark.parametrize("D", [1, 2, 6, 21, 43, 63, 80])
def test_w_inv_matches_oracle(K, fp, D):
    rng = np.random.default_rng(D)
    N, M = 37, 300
    counts = rng.integers(0, 90, size=N)
    counts[3] = 0
    counts[10] = 700  # hub row, several staging chunks
    rowptr = np.concatenate([[0], np.cumsum(counts)])
    col = rng.integers(0, M, size=rowptr[-1])
    val = rng.random(rowptr[-1])
    Xn, Xm = rng.standard_normal((N, D)), rng.standard_normal((M, D))
    ref = O.calc_W_inv(rowptr, col, val, Xn, Xm, dt=np.float64)
    for idt in (torch.int32, torch.int64):
        out = K.calc_W_inv(_t(rowptr, idt), _t(col, idt), _t(val, fp), _t(Xn, fp), _t(Xm, fp))
        tol = 2e-5 if fp == torch.float32 else 1e-12
        scale = np.abs(ref).max(axis=(1, 2), keepdims=True) + 1e-30
        assert np.max(np.abs(_np(out) - ref) / scale) < tol
        np.testing.assert_array_equal(_np(out), np.swapaxes(_np(out), 1, 2))
    assert np.all(_np(out)[3] == 0)


@pytest.mark.parametrize("D", [2, 6, 43, 63])
def test_w_inv_float16(K, D):
    """float16 inputs (csrc/cuda/W_inv_cuda.cu:102-103 dispatches kHalf):
    read into float32 arithmetic, the result rounded once to float16, so
    each entry is within half an ulp of float16 (2^-11 relative) of the
    float64 sum over the same float16 inputs, plus the float32 accumulation
    error (1e-5 of the row's largest entry)."""
    rng = np.random.default_rng(16 + D)
    N, M = 29, 200
    counts = rng.integers(0, 60, size=N)
    counts[2] = 0
    rowptr = np.concatenate([[0], np.cumsum(counts)])
    col = rng.integers(0, M, size=rowptr[-1])
    val = rng.random(rowptr[-1]).astype(np.float16)
    Xn = rng.standard_normal((N, D)).astype(np.float16)
    Xm = rng.standard_normal((M, D)).astype(np.float16)
    ref = O.calc_W_inv(rowptr, col, val.astype(np.float64), Xn.astype(np.float64),
                       Xm.astype(np.float64), dt=np.float64)
    f16 = torch.float16
    out = K.calc_W_inv(_t(rowptr, torch.int32), _t(col, torch.int64), _t(val, f16), _t(Xn, f16),
                       _t(Xm, f16))
    assert out.dtype == f16
    got = _np(out).astype(np.float64)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True) + 1e-30
    assert np.all(np.abs(got - ref) <= 2.0 ** -11 * np.abs(ref) + 1e-5 * scale)
    assert np.all(got[2] == 0)
    # the registered op writes a float16 out in place
    o2 = torch.empty_like(out)
    r = torch.ops.lakde.W_inv_cuda(_t(rowptr), _t(col), _t(val, f16), _t(Xn, f16), _t(Xm, f16), o2)
    assert r.data_ptr() == o2.data_ptr()
    np.testing.assert_array_equal(_np(o2), _np(out))


def test_w_inv_golden(K, golden):
    g = golden("ops")
    out = K.calc_W_inv(_t(g["winv_rowptr"]), _t(g["winv_col"]), _t(g["winv_val"]),
                       _t(g["winv_Xn"]), _t(g["winv_Xm"]))
    np.testing.assert_allclose(_np(out), g["winv_S"], rtol=1e-5, atol=1e-6)


# ------------------------------------------------------ chol_quad_form
def _factors(rng, N, D):
    A = rng.standard_normal((N, D, D))
    Winv = A @ np.swapaxes(A, 1, 2) + D * np.eye(D)
    L = np.linalg.cholesky(Winv)
    Linv = np.linalg.inv(L)
    return np.swapaxes(Linv, 1, 2)  # W_triu (upper)


@pytest.mark.parametrize("fp", [torch.float32, torch.float64])
@pytest.mark.parametrize("D", [1, 2, 5, 6, 9, 17, 33, 63, 70])
def test_chol_quad_form_matches_oracle(K, fp, D):
    rng = np.random.default_rng(100 + D)
    N, M = 23, 517
    Wt = _factors(rng, N, D)
    Xn, Xm = rng.standard_normal((N, D)), rng.standard_normal((M, D))
    ref = O.chol_quad_form(Wt, Xn, Xm, dt=np.float64)
    # the models pass a transposed view of the lower inverse factor
    W_dev = _t(np.swapaxes(Wt, 1, 2), fp).mT
    assert D == 1 or not W_dev.is_contiguous()
    Q = _np(K.chol_quad_form(W_dev, _t(Xn, fp), _t(Xm, fp)))
    tol = 3e-5 if fp == torch.float32 else 1e-12
    np.testing.assert_allclose(Q, ref, rtol=tol, atol=tol * ref.mean())


@pytest.mark.parametrize("D", [1, 5, 17, 63])
def test_chol_quad_form_float16(K, D):
    """float16 inputs (csrc/cuda/chol_quad_form.cu:105-106 dispatches kHalf):
    float32 arithmetic, the result rounded once to float16, against float64
    over the same float16 inputs: 2^-11 relative plus the float32 error."""
    rng = np.random.default_rng(300 + D)
    N, M = 11, 300
    Wt = _factors(rng, N, D).astype(np.float16)
    Xn = rng.standard_normal((N, D)).astype(np.float16)
    Xm = rng.standard_normal((M, D)).astype(np.float16)
    ref = O.chol_quad_form(Wt.astype(np.float64), Xn.astype(np.float64), Xm.astype(np.float64),
                           dt=np.float64)
    f16 = torch.float16
    Q = K.chol_quad_form(_t(np.swapaxes(Wt, 1, 2), f16).mT, _t(Xn, f16), _t(Xm, f16))
    assert Q.dtype == f16
    got = _np(Q).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 2.0 ** -11 * np.abs(ref) + 3e-5 * ref.mean())


def test_chol_quad_form_stride0_and_views(K):
    rng = np.random.default_rng(3)
    D, N, M = 4, 9, 50
    f32 = torch.float32
    Wt = _factors(rng, 1, D)[0]
    Xn, Xm = rng.standard_normal((N, D)), rng.standard_normal((M, D))
    W_dev = _t(Wt, f32).expand(N, D, D)       # stride-0 batch, as the shared models pass it
    Xm_view = torch.as_tensor(Xm.T.copy(), device=DEV, dtype=f32).T   # column-major points
    assert Xm_view.stride() == (1, M)
    Q1 = _np(K.chol_quad_form(W_dev, _t(Xn, f32), Xm_view))
    Q2 = _np(K.chol_quad_form(W_dev.contiguous(), _t(Xn, f32), Xm_view.contiguous()))
    np.testing.assert_array_equal(Q1, Q2)
    ref = O.chol_quad_form(np.broadcast_to(Wt, (N, D, D)), Xn, Xm, dt=np.float64)
    np.testing.assert_allclose(Q1, ref, rtol=3e-5)
    # identity factor -> squared distances; Q = 0 on coincident points
    I = torch.eye(D, device=DEV).expand(N, D, D)
    Q = _np(K.chol_quad_form(I, _t(Xn, f32), _t(Xn, f32)))
    np.testing.assert_allclose(Q, np.sum((Xn[None] - Xn[:, None]) ** 2, -1), rtol=1e-5, atol=1e-6)
    assert np.all(np.diagonal(Q) == 0)


def test_chol_quad_form_golden(K, golden):
    g = golden("ops")
    for D in (1, 2, 5, 9):
        Q = K.chol_quad_form(_t(g[f"qf{D}_Wtriu"]), _t(g[f"qf{D}_Xn"]), _t(g[f"qf{D}_Xm"]))
        np.testing.assert_allclose(_np(Q), g[f"qf{D}_Q"], rtol=2e-5, atol=1e-6)


def test_torch_ops_registered(K):
    rng = np.random.default_rng(4)
    r = _t(np.sort(rng.integers(0, 9, 30)), torch.int32)
    out = torch.empty(10, dtype=torch.int32, device=DEV)
    torch.ops.lakde.ind2ptr_cuda(r, 9, out)
    np.testing.assert_array_equal(_np(out), O.ind2ptr(_np(r), 9))
    Wt = _t(_factors(rng, 3, 2))
    Xn, Xm = _t(rng.standard_normal((3, 2))), _t(rng.standard_normal((5, 2)))
    q = torch.empty(3, 5, device=DEV, dtype=torch.float64)
    torch.ops.lakde.chol_quad_form_cuda(Wt, Xn, Xm, q)
    np.testing.assert_allclose(_np(q), _np(K.chol_quad_form(Wt, Xn, Xm)))


def test_w_inv_op_writes_out_in_place(K, golden):
    """torch.ops.lakde.W_inv_cuda writes the caller's ``out`` itself and
    returns it (csrc/cuda/W_inv_cuda.h:5-7): no new tensor is copied in."""
    g = golden("ops")
    N, D = g["winv_Xn"].shape
    out = torch.full((N, D, D), float("nan"), device=DEV)
    addr = out.data_ptr()
    ret = torch.ops.lakde.W_inv_cuda(_t(g["winv_rowptr"]), _t(g["winv_col"]), _t(g["winv_val"]),
                                     _t(g["winv_Xn"]), _t(g["winv_Xm"]), out)
    assert ret.data_ptr() == addr and out.data_ptr() == addr
    np.testing.assert_allclose(_np(out), g["winv_S"], rtol=1e-5, atol=1e-6)
    with pytest.raises((ValueError, RuntimeError)):
        torch.ops.lakde.W_inv_cuda(_t(g["winv_rowptr"]), _t(g["winv_col"]), _t(g["winv_val"]),
                                   _t(g["winv_Xn"]), _t(g["winv_Xm"]),
                                   torch.empty(N, D, D + 1, device=DEV))


# ------------------------------------------------------- fused pieces
def test_knn_prior_cov_matches_oracle():
    from lakde_amd import _fused as F
    g = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                           "golden", "ops.npz"))
    X = g["knn_X"]
    graph = g["knn_graph"]
    out = torch.empty(20, 3, 3, device=DEV)
    F.knn_prior_cov(_t(X), 0, _t(graph[:20]), out)
    np.testing.assert_allclose(_np(out), g["knn_cov"], rtol=1e-5, atol=1e-6)
    out2 = torch.empty(10, 3, 3, device=DEV)
    F.knn_prior_cov(_t(X), 20, _t(graph[20:30], torch.int64), out2)
    np.testing.assert_allclose(_np(out2), O.batched_knn_cov(X, X[20:30], graph[20:30],
                                                           np.float32), rtol=1e-5, atol=1e-6)


def _run_posterior(prior, S, nu, a, b, nu0, Xn, jit_w, jit_p=None, shared=False):
    from lakde_amd import _fused as F
    from lakde_amd._lib import PRIOR_KNN, PRIOR_SHARED
    nb, D = Xn.shape
    R = F.record_floats(D)
    outs = dict(Linv=torch.empty(nb, D, D, device=DEV), logdet=torch.empty(nb, device=DEV),
                trace=torch.empty(nb, device=DEV), rec=torch.zeros(nb, R, device=DEV),
                status=torch.zeros(1, dtype=torch.int32, device=DEV))
    jw = _t(jit_w)
    if shared:
        F.posterior(PRIOR_SHARED, _t(prior), None, _t(S), _t(nu), _t(a), _t(b), nu0, jw, _t(Xn),
                    outs["Linv"], None, outs["logdet"], None, outs["trace"], outs["rec"],
                    outs["status"])
    else:
        jp = _t(jit_p)
        outs["ptril"] = torch.empty(nb, D, D, device=DEV)
        outs["ldp"] = torch.empty(nb, device=DEV)
        F.posterior(PRIOR_KNN, _t(prior), jp, _t(S), _t(nu), _t(a), _t(b), nu0, jw, _t(Xn),
                    outs["Linv"], outs["ptril"], outs["logdet"], outs["ldp"], outs["trace"],
                    outs["rec"], outs["status"])
        outs["jit_p"] = jp
    outs["jit_w"] = jw
    return {k: _np(v) for k, v in outs.items()}


def _allowed_jitters():
    eps = np.float32(np.finfo(np.float32).eps)
    j, out = np.float32(0), {0.0}
    for k in range(32):
        j = np.float32(j + np.float32(eps * 2 ** k))
        out.add(float(j))
    return out


def test_posterior_kernel_matches_oracle():
    """Factorisation, jitter schedule, inverse, logdet, trace and the record
    constants of the posterior kernel against the oracle's float32 path.

    Rank-deficient priors (rank 3 of 6) need jitter; whether eps or 3 eps
    suffices for such a matrix depends on the rounding of the elimination, so
    there the check is that the kernel's jitter is on the reference schedule
    and factors the jittered matrix; the posterior (full rank) decisions are
    compared exactly."""
    rng = np.random.default_rng(11)
    nb, D = 40, 6
    A = rng.standard_normal((nb, D, 3)).astype(np.float32)
    prior = (A @ np.swapaxes(A, 1, 2)).astype(np.float32)      # rank 3: prior needs jitter
    prior[:10] += np.eye(D, dtype=np.float32)                   # a few full rank
    B = rng.standard_normal((nb, D, D)).astype(np.float32)
    S = (B @ np.swapaxes(B, 1, 2) * 0.1).astype(np.float32)
    S[5] = 0.01 * np.eye(D, dtype=np.float32)
    nu = (D + rng.random(nb) * 30).astype(np.float32)
    a = np.full(nb, 3.0, np.float32)
    b = (1 + rng.random(nb)).astype(np.float32)
    Xn = rng.standard_normal((nb, D)).astype(np.float32)
    jp0 = np.zeros(nb, np.float32)
    jw0 = np.zeros(nb, np.float32)
    nu0 = np.float32(D)
    out = _run_posterior(prior, S, nu, a, b, nu0, Xn, jw0, jp0)
    assert out["status"][0] == 0
    allowed = _allowed_jitters()
    jp = out["jit_p"]
    assert all(float(v) in allowed for v in jp)
    assert np.all(jp[:10] == 0) and np.all(jp[10:] > 0)
    Lp = out["ptril"].astype(np.float64)
    prior_j = O._add_diag(prior, jp)
    np.testing.assert_allclose(Lp @ np.swapaxes(Lp, 1, 2), prior_j, atol=2e-5)
    scale = (nu0 * (a / b).astype(np.float32)).astype(np.float32)
    Winv = (prior_j * scale[:, None, None]).astype(np.float32) + S
    _, jw = O.stabilized_cholesky(Winv, jw0.copy())
    np.testing.assert_array_equal(out["jit_w"], jw)
    # compare in float64 through the matrices the factors represent
    Winv_j = O._add_diag(Winv.astype(np.float64), jw.astype(np.float64))
    Linv = out["Linv"].astype(np.float64)
    W_dev = np.swapaxes(Linv, 1, 2) @ Linv          # = W (inverse of W_inv + jI)
    W_ref = np.linalg.inv(Winv_j)
    rel = np.abs(W_dev - W_ref).max(axis=(1, 2)) / np.abs(W_ref).max(axis=(1, 2))
    assert rel.max() < 1e-3
    logdet_ref = -np.linalg.slogdet(Winv_j)[1]
    np.testing.assert_allclose(out["logdet"], logdet_ref, rtol=1e-4, atol=1e-3)
    # prior logdet: exact against slogdet where the prior is well conditioned;
    # for the jittered rank-deficient ones (pivots ~ sqrt(eps)) against the
    # kernel's own factor
    np.testing.assert_allclose(out["ldp"][:10],
                               np.linalg.slogdet(prior_j[:10].astype(np.float64))[1],
                               rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(out["ldp"], 2 * np.log(np.diagonal(Lp, axis1=1, axis2=2)).sum(1),
                               rtol=1e-5, atol=1e-5)
    tr_ref = nu * O.chol_mm_trace(Lp, np.swapaxes(Linv, 1, 2))
    np.testing.assert_allclose(out["trace"], tr_ref, rtol=1e-4)
    # record constants
    T = D * (D + 1) // 2
    rec = out["rec"]
    np.testing.assert_array_equal(rec[:, T:T + D], Xn)
    tri = np.tril_indices(D)
    np.testing.assert_array_equal(rec[:, :T], out["Linv"][:, tri[0], tri[1]])
    eld = O.mvdigamma(nu / 2, D) + np.float32(D * math.log(2)) + out["logdet"]
    np.testing.assert_allclose(rec[:, T + D], 0.5 * eld, rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(rec[:, T + D + 1], -0.5 * nu)
    pc = (O.lgamma((nu + 1) / 2) - O.lgamma((nu + 1 - D) / 2) + 0.5 * out["logdet"]
          - np.float32(D / 2 * math.log(math.pi)))
    np.testing.assert_allclose(rec[:, T + D + 2], pc, rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(rec[:, T + D + 3], (nu + 1) / 2)
    # persistent jitter: a second call with the stored jitters keeps them
    out2 = _run_posterior(prior, S, nu, a, b, nu0, Xn, out["jit_w"], out["jit_p"])
    np.testing.assert_array_equal(out2["jit_w"], out["jit_w"])
    np.testing.assert_array_equal(out2["jit_p"], out["jit_p"])


def test_posterior_shared_prior():
    """Shared-prior mode: one D x D prior, trace against the full matrix
    (lakde/local_hierarchical_kde.py:289-291)."""
    rng = np.random.default_rng(12)
    nb, D = 17, 9
    C = rng.standard_normal((D, D)).astype(np.float32)
    prior = (C @ C.T + np.eye(D, dtype=np.float32)).astype(np.float32)
    B = rng.standard_normal((nb, D, D)).astype(np.float32)
    S = (B @ np.swapaxes(B, 1, 2)).astype(np.float32)
    nu = (D + rng.random(nb) * 5).astype(np.float32)
    ab = np.ones(nb, np.float32)
    Xn = rng.standard_normal((nb, D)).astype(np.float32)
    out = _run_posterior(prior, S, nu, ab, ab, np.float32(D), Xn, np.zeros(nb, np.float32),
                         shared=True)
    Winv = (prior[None] * np.float32(D)).astype(np.float32) + S
    Linv = out["Linv"].astype(np.float64)
    np.testing.assert_allclose(np.swapaxes(Linv, 1, 2) @ Linv, np.linalg.inv(Winv.astype(np.float64)),
                               rtol=1e-3, atol=1e-6)
    tr_ref = nu * O.chol_mm_trace(prior.astype(np.float64), np.swapaxes(Linv, 1, 2))
    np.testing.assert_allclose(out["trace"], tr_ref, rtol=1e-4)


@pytest.mark.parametrize("D", [20, 43, 63])
@pytest.mark.parametrize("shared", [False, True])
@pytest.mark.parametrize("jit", [0.0, 1e-3])
def test_posterior_kernel_large_d(D, shared, jit):
    """The posterior kernel at the dimensions of the d=21/43/63 workloads
    (even and odd D give different LDS row strides): W = Linv^T Linv against
    the float64 inverse of W_inv (+ the stored jitter: a nonzero one takes
    the float64 refactorisation), logdet against slogdet, the trace against
    chol_mm_trace in float64 (triangular prior factor or dense shared prior)."""
    rng = np.random.default_rng(100 + D)
    nb = 70
    nu0 = np.float32(D)
    B = rng.standard_normal((nb, D, D)).astype(np.float32)
    S = (B @ np.swapaxes(B, 1, 2) / D).astype(np.float32)
    nu = (D + rng.random(nb) * 50).astype(np.float32)
    a = np.full(nb, 2.0, np.float32)
    b = (1 + rng.random(nb)).astype(np.float32)
    Xn = rng.standard_normal((nb, D)).astype(np.float32)
    jw0 = np.full(nb, jit, np.float32)
    if shared:
        C = rng.standard_normal((D, D)).astype(np.float32)
        prior = (C @ C.T / D + np.eye(D, dtype=np.float32)).astype(np.float32)
        out = _run_posterior(prior, S, nu, a, b, nu0, Xn, jw0, shared=True)
        prior_j = np.broadcast_to(prior, (nb, D, D))
        T = prior_j.astype(np.float64)
    else:
        A = rng.standard_normal((nb, D, D)).astype(np.float32)
        prior = (A @ np.swapaxes(A, 1, 2) / D + 0.1 * np.eye(D, dtype=np.float32)).astype(np.float32)
        out = _run_posterior(prior, S, nu, a, b, nu0, Xn, jw0, np.zeros(nb, np.float32))
        assert np.all(out["jit_p"] == 0)
        prior_j = prior
        Lp = out["ptril"].astype(np.float64)
        np.testing.assert_allclose(Lp @ np.swapaxes(Lp, 1, 2), prior, rtol=1e-4, atol=1e-4)
        assert np.all(np.triu(out["ptril"], 1) == 0)
        T = Lp
    assert out["status"][0] == 0
    np.testing.assert_array_equal(out["jit_w"], jw0)
    scale = (nu0 * (a / b).astype(np.float32)).astype(np.float32)
    Winv = (prior_j * scale[:, None, None]).astype(np.float32) + S
    Winv = O._add_diag(Winv, jw0)
    Linv = out["Linv"].astype(np.float64)
    assert np.all(np.triu(out["Linv"], 1) == 0)
    W_ref = np.linalg.inv(Winv.astype(np.float64))
    W_dev = np.swapaxes(Linv, 1, 2) @ Linv
    rel = np.abs(W_dev - W_ref).max(axis=(1, 2)) / np.abs(W_ref).max(axis=(1, 2))
    assert rel.max() < 1e-3
    np.testing.assert_allclose(out["logdet"], -np.linalg.slogdet(Winv.astype(np.float64))[1],
                               rtol=1e-4, atol=1e-3)
    tr_ref = nu * O.chol_mm_trace(T, np.swapaxes(Linv, 1, 2))
    np.testing.assert_allclose(out["trace"], tr_ref, rtol=1e-4)
    tri = np.tril_indices(D)
    np.testing.assert_array_equal(out["rec"][:, :D * (D + 1) // 2], out["Linv"][:, tri[0], tri[1]])


def test_posterior_jitter_schedule_exact():
    """Rounding-independent jitter decisions: exact zero pivot -> eps;
    a -3 eps pivot -> 7 eps; a -1e-3 pivot -> (2^14 - 1) eps."""
    D, nb = 3, 3
    eps = np.finfo(np.float32).eps
    prior = np.zeros((nb, D, D), np.float32)
    prior[:] = np.eye(D, dtype=np.float32)
    S = np.zeros((nb, D, D), np.float32)
    Winv_diag = np.array([[1, 0, 2], [1, -3 * eps, 2], [1, 1, -1e-3]], np.float32)
    # W_inv = prior * scale + S with scale = 1 (nu0 = 1, a = b) -> set S = diag - I
    for i in range(nb):
        S[i] = np.diag(Winv_diag[i]) - np.eye(D, dtype=np.float32)
    ones = np.ones(nb, np.float32)
    out = _run_posterior(prior, S, ones, ones, ones, np.float32(1.0),
                         np.zeros((nb, D), np.float32), np.zeros(nb, np.float32),
                         np.zeros(nb, np.float32))
    np.testing.assert_array_equal(out["jit_p"], 0)
    np.testing.assert_array_equal(out["jit_w"], np.array([eps, 7 * eps, (2 ** 14 - 1) * eps],
                                                         np.float32))


# ----------------------------------------- one-call block posterior (fused D <= 6)
def _block_inputs(rng, N, D, nb, K, quantise):
    X = rng.standard_normal((N, D)).astype(np.float32)
    if quantise:  # repeated coordinates: rank-deficient kNN covariances, jittered
        X[:, 0] = np.round(X[:, 0] * 2) / 2
        X[:, -1] = 0.25
    graph = rng.integers(0, N, size=(nb, K)).astype(np.int32)
    if quantise:
        graph[:5] = graph[:5, :1]  # every neighbour the same point: zero covariance
    # ragged CSR rows: empty, short, and longer than one wavefront
    lens = rng.integers(0, 90, size=nb)
    lens[0], lens[1], lens[2] = 0, 1, 1500  # 1500: several passes of every wave
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(N, size=l, replace=False)) for l in lens])
    val = rng.random(col.size).astype(np.float32) * 0.5
    colscale = (0.5 + rng.random(N)).astype(np.float32)
    return X, graph, rowptr, col.astype(np.int32), val, colscale


def _block_posterior_both(X, graph, rowptr, col, val, colscale, row0, shared_prior=None):
    """lakde_block_posterior and the three separate entry points on the same
    inputs; returns both output dicts."""
    from lakde_amd import _fused as F
    from lakde_amd._lib import PRIOR_KNN, PRIOR_SHARED
    N, D = X.shape
    nb = rowptr.size - 1
    rng = np.random.default_rng(5)
    a = np.full(nb, 2.0, np.float32)
    b = (1 + rng.random(nb)).astype(np.float32)
    nu0 = np.float32(D)
    # the argument structs hold raw pointers: keep every input tensor alive
    Xd, a_t, b_t = _t(X), _t(a), _t(b)
    rp_t, col_t, val_t = _t(rowptr), _t(col), _t(val)
    cs = _t(colscale) if colscale is not None else None
    g_t = _t(graph) if graph is not None else None
    sp_t = _t(shared_prior) if shared_prior is not None else None
    outs = []
    for fused in (True, False):
        o = dict(prior=torch.empty(nb, D, D, device=DEV), S=torch.empty(nb, D, D, device=DEV),
                 nu=torch.empty(nb, device=DEV), b_out=torch.empty(nb, device=DEV),
                 jw=torch.zeros(nb, device=DEV), jp=torch.zeros(nb, device=DEV),
                 Linv=torch.empty(nb, D, D, device=DEV), ptril=torch.empty(nb, D, D, device=DEV),
                 logdet=torch.empty(nb, device=DEV), ldp=torch.empty(nb, device=DEV),
                 trace=torch.empty(nb, device=DEV),
                 rec=torch.zeros(nb, F.record_floats(D), device=DEV),
                 status=torch.zeros(1, dtype=torch.int32, device=DEV))
        if shared_prior is None:
            mode, prior, jp, ptril, ldp, g = (PRIOR_KNN, o["prior"], o["jp"], o["ptril"],
                                             o["ldp"], g_t)
        else:
            mode, prior, jp, ptril, ldp, g = PRIOR_SHARED, sp_t, None, None, None, None
        post = F.posterior_args(mode, prior, jp, o["S"], o["nu"], a_t, b_t, nu0, o["jw"],
                                Xd[row0:row0 + nb], o["Linv"], ptril, o["logdet"], ldp,
                                o["trace"], o["rec"], o["status"], b0=1e-6, b_out=o["b_out"])
        if fused:
            F.block_posterior(Xd, row0, rp_t, col_t, val_t, cs, g, post)
        else:
            if g is not None:
                F.knn_prior_cov(Xd, row0, g, prior)
            F.posterior_contribs(rp_t, col_t, val_t, Xd[row0:row0 + nb], Xd, nu0,
                                 o["S"], o["nu"], colscale=cs)
            from lakde_amd import _lib
            _lib.check(_lib.load().lakde_posterior(post, _lib.stream_of(Xd)), "posterior")
        torch.cuda.synchronize()
        outs.append({k: _np(v) for k, v in o.items()})
    return outs


@pytest.mark.parametrize("D,quantise,scaled", [(1, False, False), (2, False, True),
                                               (5, True, True), (6, False, True),
                                               (6, True, False), (7, False, True)])
def test_block_posterior_matches_separate_kernels(D, quantise, scaled):
    """lakde_block_posterior (one fused kernel for D <= 6; the separate
    kernels for D = 7) against lakde_knn_prior_cov + lakde_posterior_contribs
    + lakde_posterior on ragged CSR rows (empty, 1 entry, 1500 entries),
    jittered priors and per-column factors: nu to 2e-6 and factors and
    constants to 2e-5 relative (the sums are reduced in a different order).  Jitter decisions are equal except where a rank-deficient
    prior sits on the FP32 decision boundary (rounding of the differently
    ordered sums decides); such rows must be few and are checked to carry a
    jitter of the reference schedule."""
    rng = np.random.default_rng(100 + D)
    X, graph, rowptr, col, val, colscale = _block_inputs(rng, 3000, D, 150, 32, quantise)
    fu, ref = _block_posterior_both(X, graph, rowptr, col, val, colscale if scaled else None,
                                    row0=700)
    assert fu["status"][0] == 0 and ref["status"][0] == 0
    np.testing.assert_allclose(fu["nu"], ref["nu"], rtol=2e-6)
    if quantise:
        assert np.any(ref["jp"] > 0)
    same = (fu["jp"] == ref["jp"]) & (fu["jw"] == ref["jw"])
    assert same.mean() >= 0.9, same.mean()
    allowed = _allowed_jitters()
    assert all(float(v) in allowed for v in np.concatenate([fu["jp"], fu["jw"]]))
    if not quantise:
        bad = np.flatnonzero(~same)
        assert same.all(), (bad, fu["jp"][bad], ref["jp"][bad], fu["jw"][bad], ref["jw"][bad],
                            fu["nu"][bad], np.diff(rowptr)[bad])
    for k in ("Linv", "rec"):
        f, r = fu[k][same], ref[k][same]
        scale = np.abs(r).max(axis=tuple(range(1, r.ndim)), keepdims=True) + 1e-30
        np.testing.assert_allclose(f / scale, r / scale, rtol=0, atol=2e-5, err_msg=k)
    for k in ("logdet", "trace", "b_out"):
        np.testing.assert_allclose(fu[k][same], ref[k][same], rtol=2e-5, atol=2e-5, err_msg=k)
    # prior factors: a jittered rank-deficient prior has pivots of order
    # sqrt(jitter) that the summation order's rounding decides, so its factor
    # is compared through L L^T (the jittered matrix both paths factor) and
    # its log-determinant only where no jitter was needed
    f, r = fu["ptril"][same].astype(np.float64), ref["ptril"][same].astype(np.float64)
    ff, rr = f @ np.swapaxes(f, 1, 2), r @ np.swapaxes(r, 1, 2)
    scale = np.abs(rr).max(axis=(1, 2), keepdims=True) + 1e-30
    np.testing.assert_allclose(ff / scale, rr / scale, rtol=0, atol=2e-5, err_msg="ptril")
    full = same & (ref["jp"] == 0)
    np.testing.assert_allclose(fu["ldp"][full], ref["ldp"][full], rtol=2e-5, atol=2e-5)


def test_block_posterior_shared_prior():
    """Shared-prior mode of lakde_block_posterior (hierarchical model) against
    the separate kernels, D = 4 (fused) and D = 9 (separate)."""
    for D in (4, 9):
        rng = np.random.default_rng(300 + D)
        X, graph, rowptr, col, val, _ = _block_inputs(rng, 2000, D, 100, 16, False)
        C = rng.standard_normal((D, D)).astype(np.float32)
        prior = (C @ C.T * 0.1 + np.eye(D, dtype=np.float32)).astype(np.float32)
        fu, ref = _block_posterior_both(X, None, rowptr, col, val, None, row0=0,
                                        shared_prior=prior)
        np.testing.assert_allclose(fu["nu"], ref["nu"], rtol=2e-6)
        np.testing.assert_array_equal(fu["jw"], ref["jw"])
        for k in ("logdet", "trace", "b_out"):
            np.testing.assert_allclose(fu[k], ref[k], rtol=2e-5, atol=2e-5, err_msg=k)
        np.testing.assert_allclose(fu["rec"], ref["rec"], rtol=2e-5, atol=2e-5)
