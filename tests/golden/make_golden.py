"""Write golden input/output vectors by running the reference lakde package
(its pure-PyTorch path, LAKDE_NO_EXT=1, float32 on the CPU).

Run from the repository root, in the development container that has the
reference checkout:  python tests/golden/make_golden.py [--full]

The reference's lakde/utils.py imports TensorBoard at module load for its
logging writer; TensorBoard is not installed here, so two empty stand-in
modules are registered for that import.  The models are built with
``logs=False``, so the writer is never used and no computation goes through
the stand-ins.

Each case writes one compressed .npz next to this script holding its inputs,
its settings and the reference outputs.  The inputs come from
lakde_amd.synthetic (seeded), with small sizes so the oracle and the GPU path
can replay them in seconds.
"""

import argparse
import os
import sys
import time
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

os.environ["LAKDE_NO_EXT"] = "1"
_tb = types.ModuleType("torch.utils.tensorboard")
_tb_summary = types.ModuleType("torch.utils.tensorboard.summary")
_tb_summary.hparams = None
_tb_writer = types.ModuleType("torch.utils.tensorboard.writer")
_tb_writer.SummaryWriter = object
_tb.summary, _tb.writer = _tb_summary, _tb_writer
sys.modules["torch.utils.tensorboard"] = _tb
sys.modules["torch.utils.tensorboard.summary"] = _tb_summary
sys.modules["torch.utils.tensorboard.writer"] = _tb_writer
sys.path.insert(0, REF)

import torch  # noqa: E402

import lakde  # noqa: E402
from lakde import kernels as ref_kernels  # noqa: E402
from lakde import linalg_util as ref_linalg  # noqa: E402
from lakde import utils as ref_utils  # noqa: E402
from lakde_amd import synthetic  # noqa: E402

torch.set_num_threads(8)


def _save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
    print("wrote", path, "%.1f KB" % (os.path.getsize(path) / 1024))


def _np(t):
    return t.detach().cpu().numpy()


def case_ops():
    g = torch.Generator().manual_seed(11)
    out = {}
    # chol_quad_form with a transposed (D^2, 1, D)-strided factor, as the models pass it
    for D in (1, 2, 5, 9):
        Nb, Mb = 7, 13
        A = torch.randn(Nb, D, D, generator=g)
        Winv = A @ A.mT + D * torch.eye(D)
        W_triu = ref_linalg.tril_inverse(torch.linalg.cholesky(Winv)).mT
        Xn = torch.randn(Nb, D, generator=g)
        Xm = torch.randn(Mb, D, generator=g)
        Q = ref_kernels.chol_quad_form(W_triu, Xn, Xm)
        out[f"qf{D}_Wtriu"] = _np(W_triu.contiguous())
        out[f"qf{D}_Xn"], out[f"qf{D}_Xm"], out[f"qf{D}_Q"] = _np(Xn), _np(Xm), _np(Q)
    # calc_W_inv on a CSR with empty rows
    N, Mx, D = 9, 40, 4
    rows = torch.sort(torch.randint(0, N, (60,), generator=g)).values
    rows = rows[rows != 3]  # row 3 empty
    col = torch.randint(0, Mx, (rows.numel(),), generator=g, dtype=torch.int64).to(torch.int32)
    val = torch.rand(rows.numel(), generator=g)
    Xn = torch.randn(N, D, generator=g)
    Xm = torch.randn(Mx, D, generator=g)
    rowptr = ref_kernels.ind2ptr(rows.to(torch.int32), N)
    S = ref_kernels.calc_W_inv(rowptr, col, val, Xn, Xm)
    out.update(winv_row=_np(rows.to(torch.int32)), winv_rowptr=_np(rowptr), winv_col=_np(col),
               winv_val=_np(val), winv_Xn=_np(Xn), winv_Xm=_np(Xm), winv_S=_np(S))
    # ind2ptr with leading/trailing/inner gaps
    r = torch.tensor([2, 2, 2, 5, 5, 6, 9, 9, 9, 9], dtype=torch.int32)
    out["i2p_row"], out["i2p_ptr"] = _np(r), _np(ref_kernels.ind2ptr(r, 12))
    # special functions and log-sum-exp helpers
    a = torch.tensor([3.5, 7.25, 40.0, 130.5, 1000.0])
    for D in (2, 6, 21, 63):
        out[f"mvdigamma{D}_a"] = _np(a + D)
        out[f"mvdigamma{D}"] = _np(ref_linalg.mvdigamma((a + D) / 2, D))
        out[f"mvlgamma{D}"] = _np(ref_linalg.mvlgamma((a + D) / 2, D))
    x = torch.tensor([0.0, -1.0, -5.0, 3.0, -40.0, -np.inf])
    y = torch.tensor([-1.0, -1.0, -10.0, 2.999, -39.0, -np.inf])
    out["lse_a"], out["lse_b"], out["logsubexp"] = _np(x), _np(y), _np(ref_linalg.logsubexp(x, y))
    # stabilized_cholesky on matrices whose factor/fail decision does not
    # depend on rounding: PD, exact zero pivots (integer entries, exact
    # elimination), zero matrix, negative pivot, scaled rank-1 of integers
    M = torch.zeros(6, 5, 5)
    M[0] = torch.randn(5, 5, generator=g) @ torch.randn(5, 5, generator=g).mT * 0 + 5 * torch.eye(5)
    M[0, 0, 1] = M[0, 1, 0] = 1.0
    M[1] = torch.diag(torch.tensor([2.0, 0.0, 3.0, 1.0, 4.0]))
    ones = torch.ones(3, 3)
    M[2, :3, :3] = ones                       # pivot 2 eliminates to exactly 0
    M[2, 3, 3] = M[2, 4, 4] = 1.0
    M[3] = 0
    M[4] = torch.diag(torch.tensor([1.0, 1.0, -3 * 2.0 ** -23, 1.0, 1.0]))
    M[5] = torch.diag(torch.tensor([1.0, 1.0, 1.0, 1.0, -1e-3]))  # needs k = 13
    jitter0 = torch.zeros(6)
    L, jit = ref_linalg.stabilized_cholesky(M.clone(), jitter0.clone())
    out.update(stab_X=_np(M), stab_L=_np(L), stab_jitter=_np(jit))
    jit_in = torch.tensor([0.0, 1e-3, 0.0, 0.0, 0.0, 0.0])  # persistent jitter reused
    L2, jit2 = ref_linalg.stabilized_cholesky(M.clone(), jit_in.clone())
    out.update(stab2_jitter_in=_np(jit_in), stab2_L=_np(L2), stab2_jitter=_np(jit2))
    # batched kNN covariance
    Xk = torch.randn(50, 3, generator=g)
    graph = ref_utils.compute_knn_graph(Xk, 6, 1024, include_self=False, dtype=torch.int32)
    covk = ref_utils.batched_knn_cov(Xk, Xk[:20], graph[:20].long(), include_self=False,
                                     self_center=True)
    out.update(knn_X=_np(Xk), knn_graph=_np(graph), knn_cov=_np(covk))
    _save("ops", **out)


def _fit_case(name, model_cls, X, Y, graph, block_size, threshold, iterations,
              store_rnm=True, **kw):
    Xt = torch.from_numpy(X)
    Yt = torch.from_numpy(Y)
    m = model_cls(k_or_knn_graph=torch.from_numpy(graph), block_size=block_size,
                  verbose=False, logs=False, **kw)
    t0 = time.time()
    # the body of AbstractKDE.fit (lakde/kdes.py:46-67) with callbacks=[],
    # unrolled so the state after init and after every sweep can be recorded
    m.init_parameters_(Xt, threshold)
    m.init_summaries_(Xt, threshold)
    traj = {"init_log_consts": _np(m.rnm_log_consts).copy(), "init_nu": _np(m.nu).copy(),
            "init_W_inv_jitter": _np(m.W_inv_jitter).copy()}
    hier = hasattr(m, "W_shared_triu")
    if hier:
        traj["init_W_shared_triu"] = _np(m.W_shared_triu).copy()
    num_blocks = (Xt.size(0) - 1) // m.block_size + 1
    lc, nus, jits, bs_, wst, wsj = [], [], [], [], [], []
    for _ in range(iterations):
        for bi in range(num_blocks):
            m.partial_step_(Xt, bi, sparsity_threshold=threshold)
            m.iter_steps += 1
        lc.append(_np(m.rnm_log_consts).copy())
        nus.append(_np(m.nu).copy())
        jits.append(_np(m.W_inv_jitter).copy())
        bs_.append(_np(m.b).copy())
        if hier:
            wst.append(_np(m.W_shared_triu).copy())
            wsj.append(float(m.W_shared_inv_jitter))
    t_fit = time.time() - t0
    traj["traj_W_inv_jitter"] = np.stack(jits) if jits else np.zeros((0, X.shape[0]), np.float32)
    traj["traj_b"] = np.stack(bs_) if bs_ else np.zeros((0, X.shape[0]), np.float32)
    if hier:
        D = X.shape[1]
        traj["traj_W_shared_triu"] = np.stack(wst) if wst else np.zeros((0, D, D))
        traj["traj_W_shared_inv_jitter"] = np.asarray(wsj, np.float64)
        print(f"  {name}: posterior jitters per sweep "
              f"{[int(np.count_nonzero(j)) for j in jits]}, shared jitter {wsj}")
    t0 = time.time()
    lp = m.log_pred_density(Xt, Yt)
    t_pred = time.time() - t0
    blocks = list(m.rnm_blocks)
    rows = np.concatenate([_np(b[0][0]).astype(np.int64) + i * m.block_size
                           for i, b in enumerate(blocks)])
    cols = np.concatenate([_np(b[0][1]) for b in blocks]).astype(np.int64)
    vals = np.concatenate([_np(b[1]) for b in blocks])
    order = np.lexsort((cols, rows))
    extra = {}
    if hasattr(m, "sigma_0_jitter"):
        extra["sigma_0_jitter"] = _np(m.sigma_0_jitter)
        extra["logdet_sigma_0"] = _np(m.logdet_sigma_0)
    if hasattr(m, "W_shared_triu"):
        extra["W_shared_triu"] = _np(m.W_shared_triu)
        extra["W_shared_inv_jitter"] = _np(m.W_shared_inv_jitter)
    _save(name, X=X, Y=Y, graph=graph, block_size=block_size, threshold=threshold,
          iterations=iterations, log_pred=_np(lp), log_consts=_np(m.rnm_log_consts),
          nu=_np(m.nu), logdet_W=_np(m.logdet_W), b=_np(m.b), W_inv_jitter=_np(m.W_inv_jitter),
          traj_log_consts=np.stack(lc) if lc else np.zeros((0, X.shape[0]), np.float32),
          traj_nu=np.stack(nus) if nus else np.zeros((0, X.shape[0]), np.float32),
          time_fit_s=t_fit, time_pred_s=t_pred, rnm_nnz=rows.size,
          rnm_colsum=np.bincount(cols, weights=vals, minlength=X.shape[0]),
          **(dict(rnm_row=rows[order].astype(np.int32), rnm_col=cols[order].astype(np.int32),
                  rnm_val=vals[order]) if store_rnm else {}),
          **traj, **extra)
    print(f"  {name}: fit {t_fit:.1f}s pred {t_pred:.1f}s mean log p {float(lp.mean()):.6f}")


def _ref_graph(X, k, include_self):
    return _np(ref_utils.compute_knn_graph(torch.from_numpy(X), k, 1024,
                                           include_self=include_self, dtype=torch.int32))


def case_knn_circle_small():
    X, Y = synthetic.make("circle8", 1000, 200)
    g = _ref_graph(X, 20, False)
    _fit_case("knn_circle8_small", lakde.LocalKNearestKDE, X, Y, g, 256, 0.0, 4)


def _mix6_small():
    X, Y = synthetic.make("mix32_d6", 2000, 300)
    # coarser quantisation of coordinate 1 so that many 64-NN sets share one level
    for A in (X, Y):
        A[:, 1] = np.round(A[:, 1] * 1.5) / 1.5
    return X, Y


def case_knn_mix6_small():
    X, Y = _mix6_small()
    g = _ref_graph(X, 64, False)
    _fit_case("knn_mix6_small", lakde.LocalKNearestKDE, X, Y, g, 512, 1 / 64, 4)


def case_knn_degenerate_small():
    # coordinate 1 collapsed onto levels 10 apart: every 32-NN set shares one
    # level, so the kNN covariances are singular and the jitter path runs
    X, Y = synthetic.make("mix32_d6", 1500, 200)
    for A in (X, Y):
        A[:, 1] = 10 * np.round(A[:, 1] / 2)
    g = _ref_graph(X, 32, False)
    _fit_case("knn_degenerate_small", lakde.LocalKNearestKDE, X, Y, g, 512, 1 / 32, 3)


def case_hier_latent_small():
    X, Y = synthetic.make("latent_d43", 1200, 150)
    g = _ref_graph(X, 32, True)
    _fit_case("hier_latent_small", lakde.LocalHierarchicalKDE, X, Y, g, 512, 1 / 32, 3)


def case_hier_mix6_small():
    X, Y = synthetic.make("mix32_d6", 1500, 200)
    g = _ref_graph(X, 32, True)
    _fit_case("hier_mix6_small", lakde.LocalHierarchicalKDE, X, Y, g, 512, 1 / 32, 4)


def case_hier_smooth_small():
    # the latent-factor data without its two integer columns: well conditioned
    X, Y = synthetic.make("latent_d43", 1000, 150)
    X, Y = [np.ascontiguousarray(np.delete(A, [3, 17], axis=1)[:, :12]) for A in (X, Y)]
    g = _ref_graph(X, 24, True)
    _fit_case("hier_smooth_small", lakde.LocalHierarchicalKDE, X, Y, g, 256, 1 / 24, 4)


def _latent_smooth(n, m, dims):
    """latent_d43 without its two integer columns, first ``dims`` columns."""
    X, Y = synthetic.make("latent_d43", n, m)
    return [np.ascontiguousarray(np.delete(A, [3, 17], axis=1)[:, :dims]) for A in (X, Y)]


def _hier_jitter_counts(X, k, threshold, block_size, iterations):
    """Posterior / shared-prior jitters the reference applies per sweep."""
    Xt = torch.from_numpy(X)
    m = lakde.LocalHierarchicalKDE(k_or_knn_graph=torch.from_numpy(_ref_graph(X, k, True)),
                                   block_size=block_size, verbose=False, logs=False)
    m.init_parameters_(Xt, threshold)
    m.init_summaries_(Xt, threshold)
    out = [int(torch.count_nonzero(m.W_inv_jitter))]
    nb = (X.shape[0] - 1) // m.block_size + 1
    for _ in range(iterations):
        for bi in range(nb):
            m.partial_step_(Xt, bi, sparsity_threshold=threshold)
        out.append(int(torch.count_nonzero(m.W_inv_jitter)))
    return out, float(m.W_shared_inv_jitter)


def probe_hier_configs():
    """How the jitter-free hierarchical fixture was chosen: the reference's
    jitter counts (after init, then per sweep) on candidate configurations."""
    for n, dims, k, t in [(1200, 41, 96, 1 / 96), (1200, 20, 64, 1 / 64), (1200, 12, 48, 1 / 48),
                          (1000, 8, 32, 1 / 32), (1500, 20, 64, 0.0), (1200, 41, 128, 0.0)]:
        X, _ = _latent_smooth(n, 10, dims)
        jit, sj = _hier_jitter_counts(X, k, t, 512, 4)
        print(f"  latent_smooth N={n} D={dims} k={k} t={t:.4g}: jitters {jit}, shared {sj}",
              flush=True)
    for name, n, k, t in [("circle8", 1000, 20, 1 / 20), ("circle8", 1000, 40, 0.0),
                          ("mix64_d21", 1200, 64, 1 / 64), ("mix128_d63", 1500, 160, 1 / 160)]:
        X, _ = synthetic.make(name, n, 10)
        jit, sj = _hier_jitter_counts(X, k, t, 512, 4)
        print(f"  {name} N={n} k={k} t={t:.4g}: jitters {jit}, shared {sj}", flush=True)
    for n, k in [(3000, 256), (3000, 64), (3000, 32), (2000, 32), (6000, 32), (1200, 32),
                 (1200, 64), (2000, 128)]:
        X, _ = synthetic.make("latent_d43", n, 10)
        try:
            jit, sj = _hier_jitter_counts(X, k, 1 / k, 512, 2)
            print(f"  latent_d43 N={n} k={k}: jitters {jit}, shared {sj}", flush=True)
        except RuntimeError as e:
            print(f"  latent_d43 N={n} k={k}: {e}", flush=True)
    rng = np.random.default_rng(12)
    for dims, n, k in [(4, 1000, 24), (10, 1200, 48), (6, 1200, 32), (6, 1200, 96),
                       (8, 1200, 64)]:
        X = rng.standard_normal((n, dims)).astype(np.float32)
        jit, sj = _hier_jitter_counts(X, k, 1 / k, 512, 4)
        print(f"  normal D={dims} N={n} k={k}: jitters {jit}, shared {sj}", flush=True)


def case_hier_nojit():
    # configurations on which the reference applies no posterior and no
    # shared-prior jitter in any sweep (probe_hier_configs: every D >= 6
    # candidate collapses some posteriors onto the jitter path by the second
    # sweep), so the whole trajectory is pinned to float32 rounding
    X, Y = synthetic.make("circle8", 1000, 200)
    g = _ref_graph(X, 20, True)
    _fit_case("hier_circle8_nojit", lakde.LocalHierarchicalKDE, X, Y, g, 512, 1 / 20, 4)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((1000, 4)).astype(np.float32)
    Y = rng.standard_normal((200, 4)).astype(np.float32)
    g = _ref_graph(X, 24, True)
    _fit_case("hier_normal4_nojit", lakde.LocalHierarchicalKDE, X, Y, g, 512, 0.0, 4)


def _state_case(name, model_cls, X, Y, graph, block_size, threshold, before, after):
    """A reference model fitted `before` sweeps and written as its own save()
    writes it (torch.save of state_dict(), lakde/kdes.py:115-135), then
    fitted `after` more sweeps: the state file (name.pt) and the
    continuation's outputs (name.npz).  LocalHierarchicalKDE keeps
    rnm_blocks in a numpy object array (lakde/local_hierarchical_kde.py:175),
    which only an unpickling loader can read back; the fixture stores the
    same (indices, values) tuples as a list, so it loads with
    torch.load(weights_only=True)."""
    Xt, Yt = torch.from_numpy(X), torch.from_numpy(Y)
    m = model_cls(k_or_knn_graph=torch.from_numpy(graph), block_size=block_size, verbose=False,
                  logs=False)
    m.fit(Xt, iterations=before, callbacks=[], sparsity_threshold=threshold)
    d = m.state_dict()
    if isinstance(d["rnm_blocks"], np.ndarray):
        d["rnm_blocks"] = list(d["rnm_blocks"])
    path = os.path.join(HERE, name + ".pt")
    torch.save(d, path)
    print("wrote", path, "%.1f KB" % (os.path.getsize(path) / 1024))
    m.fit(Xt, iterations=after, callbacks=[], sparsity_threshold=threshold)
    extra = {}
    if hasattr(m, "W_shared_triu"):
        extra["W_shared_triu"] = _np(m.W_shared_triu)
    _save(name, X=X, Y=Y, graph=graph, block_size=block_size, threshold=threshold,
          before=before, after=after, log_pred=_np(m.log_pred_density(Xt, Yt)),
          log_consts=_np(m.rnm_log_consts), nu=_np(m.nu), b=_np(m.b),
          W_inv_jitter=_np(m.W_inv_jitter), **extra)


def case_state_interop():
    X, Y = _mix6_small()
    g = _ref_graph(X, 64, False)
    _state_case("state_knn_mix6", lakde.LocalKNearestKDE, X, Y, g, 512, 1 / 64, 2, 2)
    X, Y = synthetic.make("circle8", 1000, 200)
    g = _ref_graph(X, 20, True)
    _state_case("state_hier_circle8", lakde.LocalHierarchicalKDE, X, Y, g, 512, 1 / 20, 2, 2)


def case_knn_circle_full():
    w = synthetic.WORKLOADS["circle8"]
    X, Y = synthetic.make("circle8")
    g = _ref_graph(X, w["k"], False)
    _fit_case("knn_circle8_full", lakde.LocalKNearestKDE, X, Y, g, 512, 0.0, w["iterations"],
              store_rnm=False)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true", help="also run the 5k-point circle fit")
    ap.add_argument("--only", default=None)
    ap.add_argument("--probe-hier", action="store_true",
                    help="print the reference's jitter counts on candidate hierarchical "
                         "configurations (writes nothing)")
    args = ap.parse_args()
    if args.probe_hier:
        probe_hier_configs()
        sys.exit(0)
    cases = [case_ops, case_knn_circle_small, case_knn_mix6_small, case_knn_degenerate_small,
             case_hier_latent_small, case_hier_mix6_small, case_hier_smooth_small,
             case_hier_nojit, case_state_interop]
    if args.full:
        cases.append(case_knn_circle_full)
    for c in cases:
        if args.only and args.only not in c.__name__:
            continue
        print(c.__name__)
        c()
