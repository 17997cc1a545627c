"""Benchmark: VI iterations/s and test points/s of LocalKNearestKDE on one
MI355X, on the full-scale d=63 workload (mix128_d63: N=1,000,000 training and
50,000 test points, D=63, k=128, threshold 1/128, block size 512), the
precision form (BF16 screening E-step with exact FP32 refinement) against the
plain FP32 path, with the dominant kernel's roofline and the reference's CPU
path beside it.

    python bench.py [--steps K] [--warmup W] [--budget-s S] [--no-fp32]
                    [--no-cpu-baseline] [--no-secondary] [--no-tuning]

A step is one VI iteration: one sweep of partial steps over all 1954 kernel
blocks (lakde/kdes.py:64-67), with the data resident in HBM.  The K timed
iterations run through fit() of the reference API, one call per iteration,
inside one timed region.  A whole sweep of this workload takes seconds, so
the number of iterations actually timed (and of warm-up iterations) is the
largest that keeps the run within --budget-s of wall time, at most K (and W);
the line reports the numbers run, with the requested ones beside them.
Beside the headline: one whole FP32 iteration of the same workload, the
mix32_d6 FP32 path, the mix64_d21 precision form and the latent_d43
differentiation form (LocalHierarchicalKDE).  Prints one JSON line.
"""

import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
HEADLINE = "mix128_d63"
SECONDARY = "mix32_d6"

# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc runs
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) scaled by what the same
# counters read for a plain copy of known size; valid for the kernel source
# whose sha256 is recorded (the line reports null for any other source).
MEASURED_TRAFFIC = {
    "kernel": "estep_screen_kernel<63, 2, true> + estep_screen_exact_kernel<63> + "
              "estep_screen_combine_kernel<63> + screen_items_kernel<63> + "
              "screen_zero_i32_kernel<63>",
    "source_sha256": "2f705943411bd73a329a2e8627012680597771080260868973decbee9af5439d",
    "fetch_bytes": 1404500000, "write_bytes": 530395000,
    "profile": "profiles/estep_screen_mix128_d63.md: FETCH_SIZE (458,912.4 + 149,992.5 + "
               "76,894.9 + 5.0) KiB x 1.99996 + WRITE_SIZE (509,897.3 + 3,028.4 + 5,023.6 + "
               "14.9) KiB x 1.0000 per E-step (copy-probe scale factors; the keep pass, which "
               "reads 1 MB of column marks, not counted)",
}


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed VI iterations after initialisation (at least 1)")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-time budget of the whole run, which caps the iterations run")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the mix32_d6 FP32 block and the mix64_d21 precision block")
    ap.add_argument("--no-tuning", action="store_true", help="skip the prior-scale gradient leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="finish with a plain device copy of this many MiB (a known-size "
                         "reference for hardware byte counters)")
    return ap.parse_args()


def log(msg):
    print(f"# [{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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
    log(f"copy probe: {mib} MiB read + {mib} MiB written (neg kernel)")


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


def host_cpu():
    """(model name, cores usable by this process; capped at 16, the host
    share of one GPU)."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model, min(len(os.sched_getaffinity(0)), 16)


def _sweep(model, X, t, label=None):
    """One VI iteration through the reference's per-step API (the body of
    lakde/kdes.py:64-67), for the warm-up and comparison sweeps; with a
    label, a progress line every 256 partial steps."""
    nb = (X.size(0) - 1) // model.block_size + 1
    for b in range(nb):
        model.partial_step_(X, b, sparsity_threshold=t)
        if label and b % 256 == 255:
            log(f"{label}: {b + 1}/{nb} partial steps")
    model.iter_steps += nb


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
    log(f"prior scale: J {J0:.5f}, dJ/dlog s {g0:.4e} (central difference {fd:.4e}, "
        f"{t_grad:.1f}s); tuned log s {ls:+.2f}: J {J:.5f}")
    return {"points": int(Y.size(0)), "J_at_s1": J0, "dJ_dlog_s": g0,
            "central_difference": fd, "rel_diff": abs(g0 - fd) / max(abs(fd), 1e-30),
            "gradient_s": round(t_grad, 2), "path": path, "tuned_log_s": ls, "J_tuned": J}


def calibrate_c1(cores):
    """The CPU port on the circle workload (N=5,000, D=2, k=20, no threshold):
    initialisation plus 2 VI iterations, projected to the 50-iteration fit
    the reference ran in 41.8 s on 8 threads (the fit time recorded with the
    reference's outputs for this workload)."""
    from lakde_amd import synthetic
    from oracle import cpu_path as C
    from oracle import lakde_np as O
    from oracle.lakde_np import compute_knn_graph
    X, _ = synthetic.make("circle8")
    graph = compute_knn_graph(X, 20, include_self=False)
    with C.reference_cpu_ops(cores):
        m = O.LocalKNearestKDE(k_or_knn_graph=graph, block_size=512)
        t0 = time.perf_counter()
        m.init_parameters_(X, 0)
        m.init_summaries_(X, 0)
        t_init = time.perf_counter() - t0
        t0 = time.perf_counter()
        nb = (X.shape[0] - 1) // 512 + 1
        for _ in range(2):
            for b in range(nb):
                m.partial_step_(X, b, 0)
        t_it = (time.perf_counter() - t0) / 2
    proj = t_init + 50 * t_it
    return {"workload": "circle8: N=5000, D=2, k=20, threshold 0, 50 VI iterations",
            "port_init_s": round(t_init, 3), "port_s_per_iteration": round(t_it, 3),
            "port_fit_s_projected": round(proj, 2), "reference_fit_s": 41.81,
            "reference_threads": 8, "port_over_reference": round(proj / 41.81, 3)}


def cpu_baseline(model, Xh, graph_h, t, n_cols=131072):
    """The reference's CPU path timed on the host cores: the oracle's model
    code (the reference's algorithm) with the reference's torch CPU ops
    (oracle/cpu_path.py), on a bounded sample of one partial step of the
    workload, from the GPU model's state at the end of the timed iterations:
    the posterior of block 0 (kNN prior, W_inv contributions, stabilised
    Cholesky), its E-step over the first ``n_cols`` columns (scaled to all
    N), and the rebalance over all blocks; x num_blocks per VI iteration."""
    import numpy as np

    from oracle import cpu_path as C
    from oracle import lakde_np as O
    cpu_model, cores = host_cpu()
    N, D = Xh.shape
    bs = model.block_size
    nblocks = (N - 1) // bs + 1
    n_cols = min(n_cols, N)
    blocks = [(ind.cpu().numpy(), val.cpu().numpy()) for ind, val in model.rnm_blocks]
    c = model.rnm_log_consts.cpu().numpy()
    with C.reference_cpu_ops(cores):
        m = O.LocalKNearestKDE(k_or_knn_graph=graph_h, block_size=bs)
        m._init_common(Xh, t)
        m.knn_graph = np.asarray(graph_h, np.int32)
        m.k = m.knn_graph.shape[1]
        for k in ("a", "b", "nu", "W_inv_jitter", "sigma_0_jitter"):
            setattr(m, k, getattr(model, k).cpu().numpy().copy())
        m.logdet_W = np.empty(N, np.float32)
        m.logdet_sigma_0 = np.empty(N, np.float32)
        m.sigma_0_lambda_mm_trace = np.empty(N, np.float32)
        m.rnm_blocks = blocks
        t0 = time.perf_counter()
        sigma_0, _ = m.calc_prior_cov(Xh, 0)
        contrib, nu = m.calc_posterior_contribs(Xh, 0)
        W_triu = m.calc_posterior(0, sigma_0, contrib)
        t_post = time.perf_counter() - t0
        ind, val = blocks[0]
        keep = ind[1] < n_cols
        m.rnm_blocks = [(ind[:, keep], val[keep])]
        m.rnm_log_consts = c[:n_cols].copy()
        t0 = time.perf_counter()
        m.partial_expectation_step(np.ascontiguousarray(Xh[:n_cols]), 0, W_triu, nu,
                                   sparsity_threshold=t)
        t_e = time.perf_counter() - t0
        m.rnm_blocks = blocks
        m.rnm_log_consts = c
        t0 = time.perf_counter()
        m.rebalance_responsibilities_(Xh, 0, c.copy())
        t_r = time.perf_counter() - t0
    t_step = t_post + t_e * N / n_cols + t_r
    log(f"cpu baseline ({cores} threads, {cpu_model}): posterior {t_post:.2f}s, E-step "
        f"{t_e:.2f}s over {n_cols} columns, rebalance {t_r:.2f}s -> {t_step:.1f}s per partial step")
    return {"value": 1.0 / (t_step * nblocks), "unit": "VI iters/sec", "cores": cores,
            "kind": "port", "cpu_model": cpu_model,
            "sample": f"one partial step of {HEADLINE} (block 0, 512 kernels), from the GPU "
                      f"model's state after the timed iterations: posterior {t_post:.2f}s, "
                      f"E-step over the first {n_cols} of {N} columns {t_e:.2f}s (scaled "
                      f"x{N / n_cols:.3f}), rebalance over all {nblocks} blocks {t_r:.2f}s; "
                      f"{t_step:.1f}s per partial step x {nblocks} partial steps per VI "
                      f"iteration; the oracle's model code with the reference's torch CPU ops "
                      f"on {cores} threads",
            "partial_step_s": round(t_step, 3)}


def screen_source_sha():
    with open(os.path.join(ROOT, "lakde_amd", "csrc", "screen.hpp"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def secondary(steps=20, warmup=3):
    """The plain FP32 path on mix32_d6 (N=200,000, D=6, k=64, threshold
    1/64): VI iterations/s through fit() and test points/s."""
    import torch

    from lakde_amd import LocalKNearestKDE, synthetic
    from lakde_amd.utils import compute_knn_graph
    w = synthetic.WORKLOADS[SECONDARY]
    Xh, Yh = synthetic.make(SECONDARY)
    N, D = Xh.shape
    X = torch.from_numpy(Xh).cuda()
    Y = torch.from_numpy(Yh).cuda()
    t = w["threshold"]
    graph = compute_knn_graph(X, w["k"], 1024, include_self=False)
    model = LocalKNearestKDE(k_or_knn_graph=graph, block_size=512)
    model.init_parameters_(X, t)
    model.init_summaries_(X, t)
    _sweep(model, X, t)
    model.fit(X, iterations=warmup - 1, sparsity_threshold=t)
    nblocks = (N - 1) // 512 + 1
    timer = KernelTimer(every=8)
    timer.reserve(steps * nblocks)
    model._estep_timer = timer
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.fit(X, iterations=steps, sparsity_threshold=t)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    model._estep_timer = None
    k_ms, k_n = timer.mean_ms()
    model.log_pred_density(X, Y[:256])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lp = model.log_pred_density(X, Y)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    flop = N * N * D * (D + 1) / nblocks
    log(f"{SECONDARY}: {el / steps * 1e3:.1f} ms per VI iteration, E-step {k_ms:.4f} ms")
    return {"workload": f"{SECONDARY}: LocalKNearestKDE fit, N={N}, D={D}, k={w['k']}, "
                        f"threshold=1/64, block_size=512, FP32",
            "steps": steps, "warmup": warmup, "vi_iters_per_s": round(steps / el, 4),
            "ms_per_step": round(el / steps * 1e3, 3), "test_pts_per_s": round(Y.size(0) / t_pred, 1),
            "mean_log_density": float(lp.mean()),
            "estep_kernel": {"kernel": "estep_main_kernel<6,UPDATE,SKIP>", "ms": round(k_ms, 5),
                             "launches_timed": k_n,
                             "effective_tflops": round(flop / (k_ms * 1e-3) / 1e12, 2),
                             "note": "effective: the reference's nb*N pairs; the kernel skips "
                                     "the pairs its box bound excludes"}}


def precision_block(name="mix64_d21", sweeps=2):
    """The precision form against the plain FP32 path on the d=21 workload
    (mix64_d21: N=300,000, D=21, k=128, threshold 1/128): BF16 and FP32 VI
    iterations/s, each model from its own initialisation (the precision
    form's), one warm-up iteration, then `sweeps` timed iterations through
    fit(); |BF16 - FP32| of the held-out log density after the warm-up
    iteration; both predictive kernels against float64 arithmetic on the
    same records (10,000 test points)."""
    import torch

    from lakde_amd import LocalKNearestKDE, synthetic
    from lakde_amd.utils import compute_knn_graph
    w = synthetic.WORKLOADS[name]
    Xh, Yh = synthetic.make(name, n_test=10000)
    N, D = Xh.shape
    X = torch.from_numpy(Xh).cuda()
    Ys = torch.from_numpy(Yh).cuda()
    t = w["threshold"]
    graph = compute_knn_graph(X, w["k"], 1024, include_self=False)
    res = {"workload": f"{name}: LocalKNearestKDE fit, N={N}, D={D}, k={w['k']}, "
                       f"threshold=1/128, block_size=512", "steps": sweeps, "warmup": 1}
    lps = {}
    for prec in ("bf16", "fp32"):
        m = LocalKNearestKDE(k_or_knn_graph=graph, block_size=512, precision="bf16")
        t0 = time.perf_counter()
        m.init_parameters_(X, t)
        m.init_summaries_(X, t)
        torch.cuda.synchronize()
        t_init = time.perf_counter() - t0
        m.precision = prec
        _sweep(m, X, t)
        lps[prec] = m.log_pred_density(X, Ys)
        lp64 = pred_fp64(m.rec, D, Ys)
        e = (lps[prec].double() - lp64).abs()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(sweeps):
            m.fit(X, iterations=1, sparsity_threshold=t)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[prec] = {"vi_iters_per_s": round(sweeps / el, 4), "init_s": round(t_init, 2),
                     "pred_err_vs_fp64_nats": {"mean": float(e.mean()), "max": float(e.max())}}
        log(f"{name} {prec}: {el / sweeps:.3f} s per VI iteration, predictive vs float64 "
            f"{float(e.mean()):.2e} nats mean")
        del m
        torch.cuda.empty_cache()
    d = (lps["bf16"] - lps["fp32"]).abs()
    res["bf16_vs_fp32_after_one_iteration_nats"] = {"mean": float(d.mean()),
                                                    "max": float(d.max()),
                                                    "points": int(Ys.size(0))}
    res["speedup_bf16_over_fp32"] = round(res["bf16"]["vi_iters_per_s"]
                                          / res["fp32"]["vi_iters_per_s"], 3)
    return res


def c4_block(N=30000, M=3000, k=256, reps=3):
    """The differentiation form on latent_d43 (N=30,000, M=3,000, D=43, k=256,
    threshold 1/256): LocalHierarchicalKDE initialisation, the held-out log
    density with its gradient w.r.t. the query points and the 43 x 43 shared
    prior factor (HIP backward and the float64 D x D chain,
    lakde/local_hierarchical_kde.py:260-271 feeding
    lakde/local_adaptive_kde.py:200-223), timed over `reps` calls after one
    warm-up call; then one VI iteration is attempted (the reference's own fit
    of this data stops with "Stabilization did not converge")."""
    import torch

    from lakde_amd import LocalHierarchicalKDE, synthetic
    from lakde_amd.autograd import log_pred_density
    Xh, Yh = synthetic.make("latent_d43", N, M)
    X = torch.from_numpy(Xh).cuda()
    Y = torch.from_numpy(Yh).cuda()
    D = X.size(1)
    t = 1 / k
    m = LocalHierarchicalKDE(k_or_knn_graph=k, block_size=512)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.init_parameters_(X, t)
    m.init_summaries_(X, t)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0

    def fwd_bwd():
        P = m.W_shared_triu.T.detach().clone().requires_grad_(True)
        Yd = Y.clone().requires_grad_(True)
        out = log_pred_density(m, X, Yd, prior_factor=P)
        out.sum().backward()
        return float(out.detach().double().mean()), P.grad, Yd.grad

    fwd_bwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        J, gP, gY = fwd_bwd()
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / reps
    with torch.no_grad():
        log_pred_density(m, X, Y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        log_pred_density(m, X, Y)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
    res = {"workload": f"latent_d43: LocalHierarchicalKDE, N={N}, M={M}, D={D}, k={k}, "
                       f"threshold=1/{k}, block_size=512, state after initialisation",
           "init_s": round(t_init, 2),
           "forward_s": round(t_fwd, 4), "forward_test_pts_per_s": round(M / t_fwd, 1),
           "grad_s": round(t_grad, 4), "grad_test_pts_per_s": round(M / t_grad, 1),
           "mean_log_density": J,
           "grad_norms": {"query_points": float(gY.norm()), "prior_factor": float(gP.norm())},
           "note": "grad: forward + backward w.r.t. the query points and the shared prior factor "
                   "(float64 D x D chain included), per call"}
    log(f"latent_d43: init {t_init:.2f}s, log density {t_fwd * 1e3:.1f} ms, with gradients "
        f"{t_grad * 1e3:.1f} ms for {M} test points")
    t0 = time.perf_counter()
    try:
        _sweep(m, X, t)
        torch.cuda.synchronize()
        res["vi_iteration_s"] = round(time.perf_counter() - t0, 3)
    except RuntimeError as e:
        res["vi_iteration"] = f"stopped like the reference's fit: {e}"
    log(f"latent_d43: VI iteration {res.get('vi_iteration_s', res.get('vi_iteration'))}")
    return res


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    t_start = time.perf_counter()
    import torch

    from lakde_amd import LocalKNearestKDE, synthetic
    from lakde_amd.utils import compute_knn_graph

    w = synthetic.WORKLOADS[HEADLINE]
    Xh, Yh = synthetic.make(HEADLINE)
    N, D = Xh.shape
    M = Yh.shape[0]
    k = w["k"]
    t = w["threshold"]
    bs = 512
    nblocks = (N - 1) // bs + 1
    dev = torch.device("cuda:0")
    X = torch.from_numpy(Xh).to(dev)
    Y = torch.from_numpy(Yh).to(dev)
    Ys = Y[:10000]
    log(f"workload {HEADLINE}: N={N} M={M} D={D} k={k} threshold=1/128 bs={bs} "
        f"({nblocks} blocks)")
    t0 = time.perf_counter()
    graph = compute_knn_graph(X, k, 1024, include_self=False)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    log(f"knn graph {t_graph:.1f}s")

    def new_model():
        m = LocalKNearestKDE(k_or_knn_graph=graph, block_size=bs, precision="bf16")
        t0 = time.perf_counter()
        m.init_parameters_(X, t)
        m.init_summaries_(X, t)
        torch.cuda.synchronize()
        return m, time.perf_counter() - t0

    model, t_init = new_model()
    log(f"init (bf16) {t_init:.1f}s, {model.arena.total_nnz()} kept responsibilities")
    # first warm-up iteration; its state is the comparison point
    t0 = time.perf_counter()
    _sweep(model, X, t)
    torch.cuda.synchronize()
    t_sweep = time.perf_counter() - t0
    log(f"warm-up iteration 1 (bf16): {t_sweep:.2f}s")
    lp_bf = model.log_pred_density(X, Ys)
    model.precision = "fp32"
    lp_bf_fit_32pred = model.log_pred_density(X, Ys)
    model.precision = "bf16"
    torch.cuda.synchronize()

    # the FP32 path: a second model from the same initialisation, one whole
    # iteration with the FP32 E-step
    fp32 = None
    if not args.no_fp32:
        log("fp32 comparison model: initialising")
        twin, t_init2 = new_model()
        log(f"fp32 comparison model: init {t_init2:.1f}s")
        twin.precision = "fp32"
        timer32 = KernelTimer(every=16)
        timer32.reserve(nblocks)
        twin._estep_timer = timer32
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _sweep(twin, X, t, label="fp32 iteration")
        torch.cuda.synchronize()
        t_fp32 = time.perf_counter() - t0
        twin._estep_timer = None
        k32, n32 = timer32.mean_ms()
        lp32 = twin.log_pred_density(X, Ys)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        twin.log_pred_density(X, Ys)
        torch.cuda.synchronize()
        t_pred32 = time.perf_counter() - t0
        d_fit = (lp_bf_fit_32pred - lp32).abs()
        d_all = (lp_bf - lp32).abs()
        fp32 = {"vi_iters_per_s": round(1.0 / t_fp32, 6), "sweep_s": round(t_fp32, 2),
                "steps": 1, "init_s": round(t_init2, 2),
                "estep_kernel": {"kernel": f"estep_main_kernel<{D},UPDATE>",
                                 "ms": round(k32, 3), "launches_timed": n32,
                                 "algorithmic_tflops": round(bs * N * D * (D + 1) / (k32 * 1e-3)
                                                             / 1e12, 2),
                                 "peak_tflops": PEAK_FP32_TFLOPS},
                "test_pts_per_s": round(Ys.size(0) / t_pred32, 1),
                "bf16_vs_fp32_after_one_iteration_nats": {
                    "fit_only": {"mean": float(d_fit.mean()), "max": float(d_fit.max())},
                    "fit_and_predictive": {"mean": float(d_all.mean()), "max": float(d_all.max())},
                    "points": int(Ys.size(0)),
                    "note": "both models from the same initialisation, one VI iteration each "
                            "(BF16 screening vs FP32 E-step); fit_only: both predicted with the "
                            "FP32 predictive kernel"}}
        log(f"fp32 iteration {t_fp32:.1f}s (E-step {k32:.2f} ms); |bf16 - fp32| after one "
            f"iteration: fit {float(d_fit.mean()):.2e} mean, fit+pred {float(d_all.mean()):.2e}")
        del twin
        torch.cuda.empty_cache()
    # both predictive kernels against float64 arithmetic on the same records
    t0 = time.perf_counter()
    lp64 = pred_fp64(model.rec, D, Ys)
    err = {}
    for name, v in (("bf16", lp_bf), ("fp32", lp_bf_fit_32pred)):
        e = (v.double() - lp64).abs()
        err[name] = {"mean": float(e.mean()), "max": float(e.max())}
    log(f"float64 predictive on {Ys.size(0)} points in {time.perf_counter() - t0:.1f}s: "
        f"bf16 {err['bf16']['mean']:.2e}, fp32 {err['fp32']['mean']:.2e} nats mean")

    # the iterations that fit the budget (a reserve for the legs after the
    # timed region: prediction, tuning, CPU baseline, secondary block)
    # (about twice the measured durations of the legs after the timed region)
    reserve = 15.0 + (0 if args.no_tuning else 40.0) + (0 if args.no_cpu_baseline else 20.0) \
        + (0 if args.no_secondary else 45.0)
    # a second warm-up iteration through fit() (the timed path; the first
    # went through the per-step API) gives the iteration time the budget uses
    warm_done, t_it = 1, t_sweep
    if args.warmup > 1:
        t0 = time.perf_counter()
        model.fit(X, iterations=1, sparsity_threshold=t)
        torch.cuda.synchronize()
        t_it = time.perf_counter() - t0
        warm_done = 2
        log(f"warm-up iteration 2 (fit): {t_it:.2f}s")
    avail = int((args.budget_s - (time.perf_counter() - t_start) - reserve) // t_it)
    steps = max(1, min(args.steps, avail))
    warm = warm_done + max(0, min(args.warmup - warm_done, avail - steps))
    log(f"{avail} iterations fit the budget: {warm} warm-up ({warm_done} done), {steps} timed "
        f"(requested {args.warmup} + {args.steps})")
    for i in range(warm_done, warm):
        model.fit(X, iterations=1, sparsity_threshold=t)
        torch.cuda.synchronize()
        log(f"warm-up iteration {i + 1} done")

    # timed: fit() of the reference API, `steps` VI iterations, one call per
    # iteration (each call ends with the host seeing the iteration done, so
    # the run reports progress as it goes)
    timer = KernelTimer(every=8)
    timer.reserve(steps * nblocks)
    model._estep_timer = timer
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        model.fit(X, iterations=1, sparsity_threshold=t)
        log(f"timed iteration {i + 1}/{steps} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    model._estep_timer = None
    ms_step = elapsed / steps * 1e3
    k_ms, k_n = timer.mean_ms()
    log(f"timed {steps} VI iterations: {ms_step / 1e3:.3f} s each; E-step main kernel "
        f"{k_ms:.4f} ms avg over {k_n} launches")

    # test points per second: the BF16 predictive on all test points
    model.log_pred_density(X, Y[:256])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lp = model.log_pred_density(X, Y)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    mean_lp = float(lp.mean())
    log(f"predictive: {M} points in {t_pred:.2f}s, mean log density {mean_lp:.4f}")

    # roofline of the dominant kernel, the screening E-step (UPDATE mode):
    # algorithmic work D(D+1) flop per (kernel, point) pair, nb*N pairs per
    # launch; executed matrix-core work 2*32*32*16 flop per MFMA, one MFMA
    # per (32-row, 16-column) tile of the padded triangular factor that is
    # not all zeros, per 32 columns
    rt_n, ks_n = (D + 31) // 32, (D + 15) // 16
    tiles = sum(1 for rt in range(rt_n) for s_ in range(ks_n) if s_ * 16 <= rt * 32 + 31)
    pairs = bs * N
    flop = pairs * D * (D + 1)
    mfma_flop = tiles * 2 * 32 * 32 * 16 * bs * (N // 32 + (N % 32 > 0))
    achieved = flop / (k_ms * 1e-3) / 1e12
    executed = mfma_flop / (k_ms * 1e-3) / 1e12
    tr = MEASURED_TRAFFIC
    traffic, tnote = None, "not measured for this kernel source"
    if tr is not None and tr["source_sha256"] == screen_source_sha():
        traffic = tr["fetch_bytes"] + tr["write_bytes"]
        tnote = tr["profile"]
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                "traffic": traffic, "traffic_note": tnote,
                "kernel": f"estep_screen_kernel<{D},UPDATE> in its split form (the kernel ranges' "
                          f"screens, the exact values of the kernels' survivor lists, the "
                          f"per-column combine and the keep pass, one event pair around the six "
                          f"launches)", "kernel_ms": round(k_ms, 4),
                "launches_timed": k_n, "pairs_per_launch": pairs,
                "achieved_executed_mfma": round(executed, 2),
                "frac_executed_mfma": round(executed / PEAK_BF16_TFLOPS, 4),
                "note": "achieved: the reference's D(D+1) flop per pair over all nb*N pairs; "
                        "executed: the BF16 MFMA flop the screen issues (FP32 exact refinement "
                        "of the surviving pairs not counted)"}

    tuning = None
    if not args.no_tuning:
        try:
            tuning = prior_scale_tuning(model, X, Y[:2000])
        except Exception as e:  # reported, never fatal to the headline line
            log(f"prior scale tuning failed: {e!r}")
            tuning = {"error": repr(e)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(model, Xh, graph.cpu().numpy(), t)
            cpu["calibration"] = calibrate_c1(cpu["cores"])
            log(f"cpu calibration: {cpu['calibration']}")
        except Exception as e:
            log(f"cpu baseline failed: {e!r}")
            cpu = {"error": repr(e)}
    del model
    torch.cuda.empty_cache()

    sec = c3 = c4 = None
    if not args.no_secondary:
        try:
            sec = secondary()
        except Exception as e:
            log(f"secondary block failed: {e!r}")
            sec = {"error": repr(e)}
        try:
            c3 = precision_block()
        except Exception as e:
            log(f"mix64_d21 precision block failed: {e!r}")
            c3 = {"error": repr(e)}
        try:
            c4 = c4_block()
        except Exception as e:
            log(f"latent_d43 differentiation block failed: {e!r}")
            c4 = {"error": repr(e)}

    value = steps / elapsed
    out = {
        "metric": "VI iters/sec & test pts/sec, BF16+refine vs FP32 at <=1e-3 nats/pt vs FP64",
        "value": round(value, 5),
        "unit": "VI iters/sec",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 screen + f32 refine",
        "data": "synthetic",
        "config": {"workload": f"{HEADLINE}: LocalKNearestKDE fit, N={N}, D={D}, k={k}, "
                               f"threshold=1/128, block_size={bs}, precision form (BF16 "
                               f"screening E-step, exact FP32 refinement)",
                   "N": N, "M": M, "D": D, "k": k, "block_size": bs},
        "test_pts_per_sec": round(M / t_pred, 1),
        "predict_s": round(t_pred, 3),
        "mean_log_density": round(mean_lp, 6),
        "init_s": round(t_init, 2),
        "knn_graph_s": round(t_graph, 2),
        "fp32": fp32,
        "speedup_bf16_over_fp32": round(value / fp32["vi_iters_per_s"], 3) if fp32 else None,
        "pred_err_vs_fp64_nats": dict(err, points=int(Ys.size(0)),
                                      note="the predictive kernels against float64 arithmetic "
                                           "on the same records, after one VI iteration"),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "prior_scale_tuning": tuning,
        "secondary": sec,
        "precision_d21": c3,
        "differentiation_d43": c4,
        "wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)


if __name__ == "__main__":
    main()
