"""Benchmark: VI iterations/s (and test points/s) of LocalKNearestKDE on one
MI355X, with the E-step kernel's roofline and the CPU oracle beside it.

    python bench.py [--steps K] [--warmup W] [--workload mix32_d6] [--block-size 512]

A step is one VI iteration: one sweep of partial steps over all kernel blocks
of the workload (lakde/kdes.py:64-67), with the data already resident in HBM.
Prints one JSON line.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the pair-skipping estep_main_kernel<6,UPDATE> on
# mix32_d6 with block size 512: rocprofv3 FETCH_SIZE and WRITE_SIZE (separate
# runs, KiB per dispatch) scaled by what they read for a plain 512 MiB
# elementwise copy (factors 1.9999 and 1.0000); summary in
# profiles/estep_mix32_d6_hbm_traffic.md
MEASURED_TRAFFIC = {("mix32_d6", 512): {"fetch": 47249802, "write": 8159539}}


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed VI iterations after initialisation (at least 1)")
    ap.add_argument("--workload", default="mix32_d6")
    ap.add_argument("--block-size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision-form", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="finish with a plain device copy of this many MiB (a known-size "
                         "reference for hardware byte counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain elementwise copy kernel over ``mib`` MiB (b = -a: reads and
    writes exactly that many bytes), run twice."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        torch.neg(a, out=b)
        torch.cuda.synchronize()
    print(f"# copy probe: {mib} MiB read + {mib} MiB written (neg kernel)", file=sys.stderr,
          flush=True)


class KernelTimer:
    """HIP-event timing of the E-step main kernel on the stream it runs on:
    the step driver records a pair of events around its launch.  Every
    ``every``-th launch is timed (an event pair on the stream adds a few
    microseconds of launch gap, so timing every launch would slow the timed
    region itself).  Events are created (and recorded once, which creates the
    HIP event) before the timed region by ``reserve``."""

    def __init__(self, every=1):
        import torch
        self.torch = torch
        self.every = every
        self.calls = 0
        self.pool = []
        self.pairs = []

    def _new(self):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def reserve(self, n):
        self.pool = [(self._new(), self._new()) for _ in range((n - 1) // self.every + 1)]
        self.pool.reverse()
        self.torch.cuda.synchronize()

    def pair(self):
        self.calls += 1
        if (self.calls - 1) % self.every:
            return None
        p = self.pool.pop() if self.pool else (self._new(), self._new())
        self.pairs.append(p)
        return p

    def mean_ms(self):
        self.torch.cuda.synchronize()
        ts = [a.elapsed_time(b) for a, b in self.pairs]
        return sum(ts) / len(ts), len(ts)


def cpu_baseline(X, graph, block_size, threshold, seconds):
    """Time the oracle's partial step on the full workload (bounded: as many
    partial steps as fit in ``seconds``) and extrapolate to a sweep."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import lakde_np as O

    with threadpool_limits(limits=1):
        m = O.LocalKNearestKDE(k_or_knn_graph=graph, block_size=block_size)
        m.init_parameters_(X, threshold)
        # state as after initialisation, without the O(N^2) init passes:
        # normalisers from one LSE pass of block 0 only (values do not change
        # the work of a partial step)
        N = X.shape[0]
        m.rnm_log_consts = np.zeros(N, np.float32)
        m.nu[:] = m.nu_0 + 1
        t0 = time.perf_counter()
        steps = 0
        nb = (N - 1) // m.block_size + 1
        while True:
            m.partial_step_(X, steps % nb, sparsity_threshold=threshold)
            steps += 1
            if time.perf_counter() - t0 > seconds:
                break
        dt = (time.perf_counter() - t0) / steps
    return {"value": 1.0 / (dt * nb), "unit": "VI iters/sec", "cores": 1, "kind": "port",
            "sample": f"{steps} partial steps of the full workload (block size {block_size}), "
                      f"{dt:.3f} s each, extrapolated to {nb} steps per VI iteration"}


def pred_fp64(rec, D, Y, kchunk=2048, pchunk=2000):
    """The predictive density evaluated in float64 arithmetic (torch on the
    GPU) from the same float32 kernel records the HIP kernels read:
    log p(y) = -ln N + LSE_n [coeff_n - h_n log1p(||Linv_n (y - x_n)||^2)]
    (lakde/local_adaptive_kde.py:209-223).  The difference to a kernel's
    output is that kernel's arithmetic error, in nats."""
    import torch
    N = rec.size(0)
    P = D * (D + 1) // 2
    ti = torch.tril_indices(D, D, device=rec.device)
    Y64 = Y.double()
    M = Y.size(0)
    m_run = torch.full((M,), float("-inf"), dtype=torch.float64, device=Y.device)
    s_run = torch.zeros(M, dtype=torch.float64, device=Y.device)
    for k0 in range(0, N, kchunk):
        r = rec[k0:k0 + kchunk].double()
        kc = r.size(0)
        L = torch.zeros(kc, D, D, dtype=torch.float64, device=Y.device)
        L[:, ti[0], ti[1]] = r[:, :P]
        Lt = L.transpose(1, 2).contiguous()
        xc, coeff, h = r[:, P:P + D], r[:, P + D + 2], r[:, P + D + 3]
        for p0 in range(0, M, pchunk):
            d = (Y64[None, p0:p0 + pchunk, :] - xc[:, None, :]).contiguous()
            z = torch.bmm(d, Lt)
            lp = coeff[:, None] - h[:, None] * torch.log1p((z * z).sum(-1))
            mx = torch.maximum(lp.max(0).values, m_run[p0:p0 + pchunk])
            s_run[p0:p0 + pchunk] = (s_run[p0:p0 + pchunk] * torch.exp(m_run[p0:p0 + pchunk] - mx)
                                     + torch.exp(lp - mx[None]).sum(0))
            m_run[p0:p0 + pchunk] = mx
    return m_run + torch.log(s_run) - math.log(N)


def prior_scale_tuning(model, X, Y, rounds=2):
    """J(s) = mean held-out log density with the kNN prior covariances scaled
    by s: the gradient dJ/dlog s from the backward pass, a central
    difference (h = 0.01 in log s) beside the first one, and `rounds` steps of
    gradient ascent in log s, each a line search along the gradient (forward
    evaluations only)."""
    import torch

    from lakde_amd.autograd import prior_scale_objective
    t0 = time.perf_counter()
    J0, g0 = prior_scale_objective(model, X, Y)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    h = 0.01
    Jp, _ = prior_scale_objective(model, X, Y, log_scale=h, grad=False)
    Jm, _ = prior_scale_objective(model, X, Y, log_scale=-h, grad=False)
    fd = (Jp - Jm) / (2 * h)
    ls, J, g = 0.0, J0, g0
    path = []
    for r in range(rounds):
        if r > 0:
            _, g = prior_scale_objective(model, X, Y, log_scale=ls)
        direction = 1.0 if g > 0 else -1.0
        best = (ls, J)
        for step in (0.2, 0.4, 0.8, 1.6):
            Js, _ = prior_scale_objective(model, X, Y, log_scale=ls + direction * step, grad=False)
            if Js > best[1]:
                best = (ls + direction * step, Js)
        path.append({"log_s": ls, "J": J, "dJ_dlog_s": g})
        if best[0] == ls:
            break
        ls, J = best
    print(f"# prior scale: J {J0:.5f}, dJ/dlog s {g0:.4e} (central difference {fd:.4e}, "
          f"{t_grad:.1f}s); tuned log s {ls:+.2f}: J {J:.5f}", file=sys.stderr, flush=True)
    return {"points": int(Y.size(0)), "J_at_s1": J0, "dJ_dlog_s": g0,
            "central_difference": fd, "rel_diff": abs(g0 - fd) / max(abs(fd), 1e-30),
            "gradient_s": round(t_grad, 2), "path": path, "tuned_log_s": ls, "J_tuned": J}


def precision_form(seconds_budget=None):
    """The precision comparison on the full-scale d=63 workload
    (mix128_d63: N=1,000,000, D=63, k=128, threshold 1/128, block size 512):
    initialise once (BF16 form), then time partial steps of the FP32 and the
    BF16 E-step from the same state and extrapolate to a VI iteration
    (num_blocks partial steps); predictive density on 10,000 held-out points
    with both forms on the same posterior; both against float64 arithmetic
    on the same records (10,000 held-out points)."""
    import torch

    from lakde_amd import LocalKNearestKDE, synthetic
    from lakde_amd.utils import compute_knn_graph

    w = synthetic.WORKLOADS["mix128_d63"]
    Xh, Yh = synthetic.make("mix128_d63", n_test=10000)
    N, D = Xh.shape
    dev = torch.device("cuda:0")
    X = torch.from_numpy(Xh).to(dev)
    Y = torch.from_numpy(Yh).to(dev)
    t = w["threshold"]
    t0 = time.perf_counter()
    graph = compute_knn_graph(X, w["k"], 1024, include_self=False)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    print(f"# precision form: mix128_d63 N={N} D={D}, knn graph {t_graph:.1f}s", file=sys.stderr,
          flush=True)
    model = LocalKNearestKDE(k_or_knn_graph=graph, block_size=512, precision="bf16")
    t0 = time.perf_counter()
    model.init_parameters_(X, t)
    model.init_summaries_(X, t)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    nblocks = (N - 1) // model.block_size + 1
    print(f"# precision form: init (bf16) {t_init:.1f}s, {nblocks} blocks", file=sys.stderr,
          flush=True)
    res = {"workload": f"mix128_d63: LocalKNearestKDE, N={N}, D={D}, k={w['k']}, "
                       f"threshold=1/128, block_size=512",
           "init_s_bf16": round(t_init, 2), "num_blocks": nblocks}
    steps = 6
    for prec in ("fp32", "bf16"):
        model.precision = prec
        timer = KernelTimer()
        timer.reserve(steps + 1)
        model._estep_timer = timer
        model.partial_step_(X, 0, sparsity_threshold=t)  # warm-up
        torch.cuda.synchronize()
        timer.pairs.clear()
        t0 = time.perf_counter()
        for i in range(steps):
            model.partial_step_(X, 1 + i * (nblocks // steps), sparsity_threshold=t)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        k_ms, _ = timer.mean_ms()
        model._estep_timer = None
        pairs = 512 * N
        res[prec] = {"partial_step_ms": round(dt * 1e3, 2),
                     "vi_iters_per_s_extrapolated": round(1.0 / (dt * nblocks), 5),
                     "estep_main_ms": round(k_ms, 3),
                     "estep_algorithmic_tflops": round(pairs * D * (D + 1) / (k_ms * 1e-3) / 1e12, 2)}
        print(f"# precision form {prec}: {dt * 1e3:.1f} ms/partial step, E-step main {k_ms:.2f} ms",
              file=sys.stderr, flush=True)
    # executed matrix-core work of the bf16 E-step main kernel: one pass of
    # 3 (bf16x3) x 2 x 32 x 16 flop per pair for every 32-row x 16-column
    # tile of the padded Linv that is not all zeros (the UPDATE step tests
    # its queued candidates without a second pass)
    rt_n, ks_n = (D + 31) // 32, (D + 15) // 16
    tiles = sum(1 for rt in range(rt_n) for s_ in range(ks_n) if s_ * 16 <= rt * 32 + 31)
    mfma_flop = 3 * 2 * 32 * 16 * tiles * 512 * N
    res["bf16"]["mfma_executed_tflops"] = round(mfma_flop / (res["bf16"]["estep_main_ms"] * 1e-3) / 1e12, 1)
    res["bf16"]["mfma_frac_of_2500_tflops"] = round(res["bf16"]["mfma_executed_tflops"] / 2500.0, 4)
    res["speedup_bf16_over_fp32"] = round(res["fp32"]["partial_step_ms"] / res["bf16"]["partial_step_ms"], 2)
    # predictive density, both forms, same posterior
    lps = {}
    for prec in ("fp32", "bf16"):
        model.precision = prec
        model.log_pred_density(X, Y[:256])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lps[prec] = model.log_pred_density(X, Y)
        torch.cuda.synchronize()
        res[prec]["test_pts_per_s"] = round(Y.size(0) / (time.perf_counter() - t0), 1)
    d = (lps["bf16"] - lps["fp32"]).abs()
    res["pred_abs_err_bf16_vs_fp32_nats"] = {"mean": float(d.mean()), "max": float(d.max())}
    res["mean_log_density_fp32"] = float(lps["fp32"].mean())
    # both forms against float64 arithmetic on the same records, on the
    # 10,000 held-out points
    t0 = time.perf_counter()
    lp64 = pred_fp64(model.rec, D, Y)
    torch.cuda.synchronize()
    for prec in ("fp32", "bf16"):
        e = (lps[prec].double() - lp64).abs()
        res[prec]["pred_err_vs_fp64_nats"] = {"mean": float(e.mean()), "max": float(e.max()),
                                              "points": int(Y.size(0))}
    print(f"# precision form: float64 predictive on {Y.size(0)} points in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # the differentiation form at full scale: tune a multiplicative scale s
    # of the kNN prior covariances by the held-out gradient dJ/dlog s
    # (HIP backward of the predictive density + float64 Cholesky VJPs over
    # all 1M posteriors), checked by central differences of J
    try:
        res["prior_scale_tuning"] = prior_scale_tuning(model, X, Y[:2000])
    except Exception as e:  # reported, never fatal to the headline line
        print(f"# prior scale tuning failed: {e!r}", file=sys.stderr, flush=True)
        res["prior_scale_tuning"] = {"error": repr(e)}
    del model, X, Y, graph
    torch.cuda.empty_cache()
    return res


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from lakde_amd import LocalKNearestKDE, synthetic
    from lakde_amd import _fused as F
    from lakde_amd import _lib
    from lakde_amd.utils import compute_knn_graph

    w = synthetic.WORKLOADS[args.workload]
    Xh, Yh = synthetic.make(args.workload)
    N, D = Xh.shape
    M = Yh.shape[0]
    dev = torch.device("cuda:0")
    X = torch.from_numpy(Xh).to(dev)
    Y = torch.from_numpy(Yh).to(dev)
    t = w["threshold"]
    print(f"# workload {args.workload}: N={N} M={M} D={D} k={w['k']} threshold={t:.6g} "
          f"bs={args.block_size}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    graph = compute_knn_graph(X, w["k"], 1024, include_self=False)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0

    model = LocalKNearestKDE(k_or_knn_graph=graph, block_size=args.block_size)
    nblocks = (N - 1) // min(args.block_size, N) + 1
    # initialisation and the warm-up iterations, untimed
    warm = max(args.warmup, 1)
    t0 = time.perf_counter()
    model.init_parameters_(X, t)
    model.init_summaries_(X, t)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    print(f"# knn graph {t_graph:.2f}s, init {t_init:.2f}s, nnz {model.arena.total_nnz()}",
          file=sys.stderr, flush=True)
    for i in range(warm):
        for b in range(nblocks):
            model.partial_step_(X, b, sparsity_threshold=t)
        model.iter_steps += nblocks
    print(f"# {warm} warm-up iteration(s) done", file=sys.stderr, flush=True)

    # timed: fit() of the reference API, K VI iterations (one native sweep
    # loop each), the responsibilities folded at the end
    timer = KernelTimer(every=8)
    timer.reserve(args.steps * nblocks)
    model._estep_timer = timer
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.fit(X, iterations=args.steps, sparsity_threshold=t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    model._estep_timer = None
    ms_step = elapsed / args.steps * 1e3
    k_ms, k_n = timer.mean_ms()
    print(f"# timed {args.steps} VI iterations: {ms_step:.1f} ms each; E-step main kernel "
          f"{k_ms:.4f} ms avg over {k_n} launches", file=sys.stderr, flush=True)

    # predictive density
    lp = model.log_pred_density(X, Y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        lp = model.log_pred_density(X, Y)
    torch.cuda.synchronize()
    t_pred = (time.perf_counter() - t0) / reps
    mean_lp = float(lp.mean())

    # roofline of the E-step main kernel (UPDATE mode): algorithmic work is
    # D(D+1) FLOP per (kernel, point) pair, nb * N pairs per launch -- the
    # reference's pairs; with pair skipping (the default for D <= 16) the
    # kernel evaluates only the pairs its bounds cannot exclude
    bs = model.block_size
    pairs_per_launch = (N * N) / nblocks
    flop = pairs_per_launch * D * (D + 1)
    achieved = flop / (k_ms * 1e-3) / 1e12
    tr = MEASURED_TRAFFIC.get((args.workload, bs))
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": (tr["fetch"] + tr["write"]) if tr else None,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)",
                "kernel": f"estep_main_kernel<{D},UPDATE{',SKIP' if model._col_order is not None else ''}>",
                "kernel_ms": round(k_ms, 5), "pairs_per_launch": int(pairs_per_launch)}

    prec = None
    if not args.no_precision_form:
        try:
            prec = precision_form()
        except Exception as e:  # reported, never fatal to the headline line
            print(f"# precision form failed: {e!r}", file=sys.stderr, flush=True)
            prec = {"error": repr(e)}

    cpu = None
    if not args.no_cpu_baseline:
        g = graph.cpu().numpy()
        cpu = cpu_baseline(Xh, g, bs, t, args.cpu_seconds)

    out = {
        "metric": "VI iters/sec & test pts/sec, BF16+refine vs FP32 at <=1e-3 nats/pt vs FP64",
        "value": round(1e3 / ms_step, 4),
        "unit": "VI iters/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: LocalKNearestKDE fit, N={N}, D={D}, "
                               f"k={w['k']}, threshold=1/{round(1 / t) if t else 0}, "
                               f"block_size={bs}, FP32",
                   "N": N, "M": M, "D": D, "k": w["k"], "block_size": bs},
        "test_pts_per_sec": round(M / t_pred, 1),
        "predict_ms": round(t_pred * 1e3, 3),
        "mean_log_density": round(mean_lp, 6),
        "init_s": round(t_init, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "precision_form": prec,
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)


if __name__ == "__main__":
    main()
