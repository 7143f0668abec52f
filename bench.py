"""Benchmark of the admixture_qn! iteration on one MI355X.

One step = one admixture_qn! iteration (mode=:ZAL, Q=3) on the device-resident
state: 2 update_q! + 2 update_p! (gradient/Hessian passes over all I x J
genotypes + batched QPs + projections), 2 log-likelihood passes, the ZAL
extrapolation and the accept rule.  Workload (default): I=100,000
individuals, J=100,000 SNPs, K=8, synthetic cohort (Q ~ Dirichlet(0.3),
P ~ Beta(0.5,0.5) clipped to [0.01,0.99], 2% missing) generated on the device
into the 2-bit BED (2.5 GB in HBM), random init + 5 FRAPPE-EM priming steps
before timing.  value = I*J*steps / elapsed seconds (genotype-iterations/s)
of the plain FP64 method.

Time to converge: from the same random start, each form (plain FP64, and the
BF16-matrix-core update passes under FP64 log-likelihood/QP/convergence) runs
EM(5) priming plus admixture_qn! iterations; its time to converge is the wall
time from the start of priming until the first iteration whose log-likelihood
is within 1e-4 (relative) of LL*, the final log-likelihood of the plain FP64
run (rtol 1e-7).  The warm-start forms (FP64 or BF16 passes) start from the
learned proposal instead (features, network inference, P refinement and the
EM judge are inside the timed span); their network is trained beforehand on
cohorts simulated from this workload's populations (reported separately).  The CPU figure multiplies the oracle's measured seconds per
EM step and per iteration (bounded sample, scaled to all SNPs) by the plain
run's step counts.

The "skfr" object times the SKFR AIM pre-selection (sparse k-means with
feature ranking, k = K clusters, J/10 AIMs) on the same resident BED, with
the roofline of its centers pass and the single-threaded oracle beside it.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FLOPs per genotype of the algebra each pass kernel executes (FMA = 2;
# multiply, add, reciprocal and log = 1), the basis of `roofline`:
#   qside_update  FP64 rank-1 form over the upper triangle: s = q.p (2K-1),
#                 t, s*t, 1/(s t), 1/s, 1/t, c, d, a, b, a+b, sum b, sum d (12),
#                 (a+b) p_k, hb_k += b p_k (3K), K(K+1)/2 FMAs of the triangle
#                 (K^2+K); the gradient is formed per individual from the
#                 triangle, sum b p and sum d after the pass   -> K^2 + 6K + 11
#   pside_update  s (2K-1), t, s*t, 1/(s t), 1/s, 1/t, c, d, a+b as c/s + d/t (3),
#                 c-d (11), (a+b) q_k, z_k += (c-d) q_k (3K), triangle (K^2+K)
#                                                              -> K^2 + 6K + 10
#   qside_ll      FP64 product form: s (2K-1), t, s^g t^(2-g) as two products (3)
#                                                              -> 2K + 2
#   *_f32         FP32 direct two-term form of the Q side: s, t, weights (2K+8),
#                 1-p_k, a p_k, b (1-p_k), z_k two FMAs (7K), two FMAs per
#                 triangle entry (2K^2+2K)                     -> 2K^2 + 11K + 8
#   *_bf16        vector part only (the Hessian runs on the matrix cores, counted
#                 apart): s, t, weights, c-d, sum d (2K+10), z_k (2K) -> 4K + 10
# The reference's own count for its full K x K two-term form (4K^2+8K+6 Q side,
# 2K^2+5K+8 P side, 2K+7 log-likelihood) is larger than what these kernels
# execute, so it is kept as the algorithmic reference figure, not the roofline.
EXEC_FLOPS = {
    "qside_update": lambda K: K * K + 6 * K + 11,
    "pside_update": lambda K: K * K + 6 * K + 10,
    "qside_ll": lambda K: 2 * K + 2,
    "qside_update_f32": lambda K: 2 * K * K + 11 * K + 8,
    "pside_update_f32": lambda K: K * K + 6 * K + 10,
    "qside_update_bf16": lambda K: 4 * K + 10,
    "pside_update_bf16": lambda K: 4 * K + 10,
}
REF_FLOPS = {
    "qside_update": lambda K: 4 * K * K + 8 * K + 6,
    "pside_update": lambda K: 2 * K * K + 5 * K + 8,
    "qside_ll": lambda K: 2 * K + 7,
}
# MI355X vector-ALU peaks in TFLOP/s: FP64 78.6; FP32 78.6 without packed FP32
# (every build here disables v_pk_*_f32, which the 157.3 TF FP32 figure needs).
PEAK_VALU_TFLOPS = {"f64": 78.6, "f32": 78.6}
PEAK_MFMA_BF16_TFLOPS = 2500.0  # MI355X dense BF16 / FP16 matrix peak


def mfma_pass_flops(kernel: str, K: int) -> int:
    """Executed matrix-core FLOPs per genotype of the BF16 form's matrix-core
    update passes (padding of the tiles included): per 32 x 32 block, s takes
    ceil(6K/16) v_mfma_f32_32x32x16 (32768 FLOPs each); the Hessian
    ceil(K(K+1)/32) 16x16x32 tiles (16384 FLOPs) per (half, group) on the Q
    side (weights a, b), one per group on the P side (weight a + b); the
    gradient, with two-term weights against two-row tables, 8 16x16x32 (K <= 8)
    or 8 32x32x16 (Q side) and half that on the P side."""
    nm = (6 * K + 15) // 16
    nt16 = (K * (K + 1) // 2 + 15) // 16
    if kernel == "qside_update_bf16":
        eq = nm + 2 * nt16 + (4 if K <= 8 else 8)
    else:
        eq = nm + nt16 + (2 if K <= 8 else 4)
    return 32 * eq  # 32768 FLOPs per equivalent 32x32x16 MFMA / 1024 genotypes

WORKLOADS = {
    "c3": dict(I=100_000, J=100_000, K=8, alpha=0.3, p_dist="beta", p_lo=0.01, p_hi=0.99, miss_rate=0.02),
    "c2": dict(I=10_000, J=100_000, K=5, alpha=0.3, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.02),
    "c1": dict(I=1_000, J=10_000, K=3, alpha=0.5, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.01),
    # learned warm start: 10-population world (Balding-Nichols, F=0.05), training
    # cohorts I=5,000 x J=20,000 (the panel SNPs), evaluation cohort of all 10
    "c5": dict(I=200_000, J=50_000, K=10, alpha=0.5, train_I=5_000, train_J=20_000, train_cohorts=14,
               miss_rate=0.01),
    # K sweep K=2..12 with 5-fold CV on one cohort of true K=6
    "c4": dict(I=20_000, J=200_000, K=6, alpha=0.5, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.01),
}


KERNEL_SYMBOLS = {  # bench kernel name -> rocprof symbol prefix of its K=8 (skipmissing) instance
    "qside_update": "void admx::k_qside<double, 8, 0, true, false",
    "pside_update": "void admx::k_pside<double, 8, 0, false",
    "qside_ll": "void admx::k_qside<double, 8, 2, false, false",
    "qside_update_bf16": "_ZN4admx12k_qside_bf16ILi8E",
    "pside_update_bf16": "void admx::k_pside_bf16<8>",
    # the BF16 form's matrix-core update passes (the default from K = 7)
    "qside_update_bf16_mfma": "void admx::k_qside_mfma<8, false>",
    "pside_update_bf16_mfma": "void admx::k_pside_mfma<8>",
}


def measured_traffic(kernel: str, workload: str, tag: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload and form (tag f64, f32 or bf16; already
    scaled by the copy calibration); None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{tag}.txt")
    sym = KERNEL_SYMBOLS.get(kernel)
    if sym is None or not os.path.exists(path):
        return None
    for line in open(path):
        parts = [x.strip() for x in line.split("|")]
        if len(parts) == 5 and parts[0].startswith(sym):
            return (float(parts[3]) + float(parts[4])) * 1e9
    return None


def host_cores() -> int:
    n = os.environ.get("OMP_NUM_THREADS")
    if n:
        return max(1, int(n))
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    """The host CPU's model name (first 'model name' of /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, dtype, seed, budget_s=20.0, full_snps=False, ctx=None):
    """The CPU oracle (C/OpenMP restatement of the reference loops) on a bounded
    sample of the same workload: all I individuals, the first Js SNPs (same
    truth, same generator), one EM priming step, then admixture_qn!
    iterations until about budget_s seconds of CPU work are spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    import __graft_entry__ as g

    oa = g.load_package()
    from importlib import import_module

    sim = import_module("openadmixture_jl_amd.sim")
    I, K = wl["I"], wl["K"]
    Js = wl["J"] if full_snps else max(64, min(wl["J"], int(2e8 // I)))
    Qt, Pt = sim.draw_truth(I, wl["J"], K, wl["alpha"], wl["p_dist"], wl["p_lo"], wl["p_hi"], seed=seed)
    Pt = np.asfortranarray(Pt[:, :Js])
    pitch = (I + 3) // 4
    if Js == wl["J"] and ctx is not None:
        # the whole cohort: the device generated the same bytes (the generators agree bit for bit)
        bed = np.ascontiguousarray(ctx.download_bed().ravel(order="F"))
    else:
        bed = O.gen_bed(Qt, Pt, seed, wl["miss_rate"], pitch)
    T = np.float64 if dtype == "f64" else np.float32
    rng = np.random.default_rng(seed + 1)
    x0 = rng.random((K, I + Js))
    x0[:, :I] /= x0[:, :I].sum(0)
    N = K * (I + Js)
    cores = host_cores()
    st = O.OracleAdmix(bed, pitch, I, Js, K, 3, x0, rng.random((N, 3)), rng.random((N, 3)), dtype=T,
                       nthreads=cores)
    t0 = time.perf_counter()
    st.init_em(1)
    em_s = time.perf_counter() - t0
    its, t0 = 0, time.perf_counter()
    while True:
        st.qn_iteration(its + 1)
        its += 1
        el = time.perf_counter() - t0
        print(f"cpu baseline: {its} iterations over {Js} SNPs in {el:.1f} s", file=sys.stderr, flush=True)
        if its >= 2 and (el > budget_s or its >= 50 or el * (its + 1) / its > 1.5 * budget_s):
            break
    return {
        "value": I * Js * its / el,
        "unit": "genotype*iter/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle (C/OpenMP restatement of src/loops.jl etc.) on {cores} threads of {cpu_model()}, "
                  f"I={I}, {'all' if Js == wl['J'] else 'first'} J={Js} SNPs of the same cohort, K={K}, {dtype}, "
                  f"{its} full admixture_qn! iterations after 1 EM step, {el:.1f} s",
        "cpu_model": cpu_model(),
        "em_step_seconds_full": em_s * wl["J"] / Js,
        "iteration_seconds_full": el / its * wl["J"] / Js,
    }


def oracle_converge(ctx, wl, seed, max_iter, tol, rtol=1e-7):
    """The CPU oracle from the plain run's start (same seeded random x, EM(5))
    to rtol on the device cohort's genotypes: its final LL, and the wall time
    until its LL is first within tol of that final value (all host threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    import __graft_entry__ as g

    oa = g.load_package()
    I, J, K = wl["I"], wl["J"], wl["K"]
    bed = np.ascontiguousarray(ctx.download_bed().ravel(order="F"))
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(seed + 1))
    st = O.OracleAdmix(bed, (I + 3) // 4, I, J, K, 3, d.x, nthreads=host_cores())
    t0 = time.perf_counter()
    st.init_em(5)
    hist = [(time.perf_counter() - t0, st.ll_new)]
    for it in range(1, max_iter + 1):
        r = st.qn_iteration(it)
        hist.append((time.perf_counter() - t0, r["new"]))
        if r["reldiff"] < rtol:
            break
    ll = hist[-1][1]
    hit = first_within(hist, ll, tol)
    return {"ll_final": ll, "iterations": len(hist) - 1, "seconds_to_tol": hit[1], "qn_iterations_to_tol": hit[0],
            "seconds_total": hist[-1][0], "cores": host_cores(), "cpu_model": cpu_model()}


def converge_run(oa, ctx, wl, T, seed, form, max_iter, rtol=1e-7, stop_at=None):
    """EM(5) priming + admixture_qn! iterations from the seeded random start with
    the update passes in `form`; returns [(seconds since start, LL)] per step
    (priming as one entry) and the number of QN iterations run.  With stop_at =
    (LL*, tol) the run ends at the first iteration within tol of LL*."""
    I, J, K = wl["I"], wl["J"], wl["K"]
    ctx.set_precision("f64" if form == "plain" else form)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(seed + 1), T=T)
    ctx.synchronize()
    t0 = time.perf_counter()
    oa.init_em(d, None, 5, d_cu=ctx, log=None)
    ctx.set_ll(d.ll_prev, d.ll_new)  # a fresh run writes every QN history column before reading it
    hist = [(time.perf_counter() - t0, d.ll_new)]
    its = 0
    for it in range(1, max_iter + 1):
        r = ctx.qn_iteration(it)
        hist.append((time.perf_counter() - t0, r["new"]))
        its = it
        if r["reldiff"] < rtol:
            break
        if stop_at is not None and abs(r["new"] - stop_at[0]) <= stop_at[1] * abs(stop_at[0]):
            break
    return hist, its


def _warm_kw(args):
    kw = {} if args.warm_p_steps is None else {"p_steps": args.warm_p_steps}
    if args.warm_q_steps is not None:
        kw["q_steps"] = args.warm_q_steps
    return kw


def first_within(hist, ll_star, tol):
    """(index, seconds, LL) of the first entry of hist within tol of ll_star."""
    for n, (t, ll) in enumerate(hist):
        if abs(ll - ll_star) <= tol * abs(ll_star):
            return n, t, ll
    return None


def run_sweep(args):
    """K = 2..12 sweep with 5-fold cross-validation on one device cohort; prints
    one JSON line with the per-K time and CV error."""
    import __graft_entry__ as g

    oa = g.load_package()
    from importlib import import_module

    sim = import_module("openadmixture_jl_amd.sim")
    wl = WORKLOADS[args.workload]
    I, J, Kt = wl["I"], wl["J"], wl["K"]
    src = oa.HIPAdmixData(I, J, Kt, 3)
    sim.cohort_on_device(src, wl["alpha"], wl["p_dist"], wl["p_lo"], wl["p_hi"], wl["miss_rate"], seed=args.seed)
    Ks = [int(k) for k in args.sweep_ks.split(",")] if args.sweep_ks else list(range(2, 13))
    t0 = time.perf_counter()
    res = oa.cv_sweep(src, Ks, nfolds=args.folds, fold_seed=args.seed, init_seed=args.seed + 1,
                      max_iter=args.sweep_max_iter, precision="f64" if args.form == "plain" else args.form,
                      fold_start=args.fold_start, log=lambda m: print(m, file=sys.stderr, flush=True),
                      workers=args.sweep_workers)
    el = time.perf_counter() - t0
    src.close()
    out = {
        "metric": "K-sweep wall time and CV error (held-out binomial deviance per entry)",
        "value": el, "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "dtype": "f64" if args.form == "plain" else f"{args.form} passes, f64 LL/QP",
        "data": "synthetic cohort generated on device (Dirichlet Q, Uniform P, Binomial(2, q.p) genotypes)",
        "config": {"workload": f"{args.workload}: I={I}, J={J}, true K={Kt}, K={Ks[0]}..{Ks[-1]}, {args.folds}-fold CV",
                   "I": I, "J": J, "K_true": Kt, "folds": args.folds, "max_iter": args.sweep_max_iter,
                   "fold_start": args.fold_start, "workers": args.sweep_workers},
        "best_k": oa.best_k(res),
        "per_k": [{k: r[k] for k in ("K", "seconds", "cv_error", "ll_full", "iterations_full", "converged_full",
                                     "iterations_folds", "waves")} for r in res],
    }
    print(json.dumps(out), flush=True)


def run_warmstart(args):
    """Learned warm start: simulate and solve the training cohorts, train the
    network on the device, then time plain / BF16 / warm / warm+BF16 starts to
    |LL - LL*| <= 1e-4 |LL*| on the evaluation cohort; prints one JSON line."""
    import __graft_entry__ as g

    oa = g.load_package()
    wl = WORKLOADS[args.workload]
    I, J, K = wl["I"], wl["J"], wl["K"]
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    P, _ = oa.world_frequencies(K, J, fst=0.05, seed=args.seed)
    panel = oa.Panel(P[:, :wl["train_J"]].mean(0), seed=args.seed)
    t0 = time.perf_counter()
    X, T, info = oa.training_set(P, panel, wl["train_cohorts"], I=wl["train_I"], J=wl["train_J"], Ks=range(4, 11),
                                 alpha_range=(0.1, 1.0), miss_rate=wl["miss_rate"], seed=args.seed, log=log)
    t_data = time.perf_counter() - t0
    net = oa.MLP(64, 256, K, seed=args.seed)
    t0 = time.perf_counter()
    hist = net.train(X, T, epochs=args.train_epochs, batch=512, lr=2e-3, seed=args.seed)
    t_train = time.perf_counter() - t0
    log(f"training set {X.shape[0]} rows in {t_data:.1f} s; network trained in {t_train:.2f} s, "
        f"loss {hist[0]:.4f} -> {hist[-1]:.4f}")

    rng = np.random.default_rng(args.seed + 99)
    Qt = np.asfortranarray(rng.dirichlet(wl["alpha"] * np.ones(K), size=I).T)
    # the evaluation cohort's populations: drifts (F = --warm-drift) of the panel's
    P_eval = P if args.warm_drift <= 0 else oa.drifted_populations(P, args.warm_drift, args.seed + 17)
    ctx = oa.HIPAdmixData(I, J, K, 3)
    ctx.generate_bed(Qt, P_eval, args.seed + 5, wl["miss_rate"])
    tol = 1e-4
    runs = {}

    def plain(form, stop_at):
        ctx.set_precision(form)
        d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(args.seed + 1))
        ctx.synchronize()
        t_start = time.perf_counter()
        oa.init_em(d, None, 5, d_cu=ctx, log=None)
        ctx.set_ll(d.ll_prev, d.ll_new)  # a fresh run writes every QN history column before reading it
        h = [(time.perf_counter() - t_start, d.ll_new)]
        for it in range(1, args.ttc_max_iter + 1):
            r = ctx.qn_iteration(it)
            h.append((time.perf_counter() - t_start, r["new"]))
            if r["reldiff"] < 1e-7:
                break
            if stop_at and abs(r["new"] - stop_at) <= tol * abs(stop_at):
                break
        return h, None, None

    def warm(form, stop_at):
        ctx.set_precision(form)
        r = oa.warm_started_fit(ctx, net, panel, seed=args.seed + 1, max_iter=args.ttc_max_iter, **_warm_kw(args),
                                stop_at=(stop_at, tol) if stop_at else None)
        return r["history"], r["chosen"], r["gain"]

    h, _, _ = plain("f64", None)
    ll_star = h[-1][1]
    runs["plain_f64"] = (h, None, None)
    runs["bf16"] = plain("bf16", ll_star)
    runs["warm_f64"] = warm("f64", ll_star)
    runs["warm_bf16"] = warm("bf16", ll_star)
    ctx.close()
    net.close()
    res = {}
    for name, (h, chosen, _) in runs.items():
        hit = first_within(h, ll_star, tol)
        res[name] = {"seconds": hit[1] if hit else None, "qn_iterations": hit[0] if hit else None,
                     "start_seconds": h[0][0], "start_ll": h[0][1], "ll_last": h[-1][1]}
        if chosen:
            res[name]["chosen_start"] = chosen
            res[name]["judge_gain"] = runs[name][2]
    base = res["plain_f64"]["seconds"]
    for name in res:
        if res[name]["seconds"]:
            res[name]["speedup_vs_plain"] = base / res[name]["seconds"]
    out = {
        "metric": "time to |LL - LL*| <= 1e-4 |LL*| with the learned warm start (LL* = plain FP64 final LL)",
        "value": res["warm_bf16"]["seconds"], "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "dtype": "bf16 passes, f64 LL/QP; f32 network",
        "data": "synthetic: 10-population Balding-Nichols world (F=0.05), Dirichlet(alpha) admixture; evaluation "
                f"cohort from drifted copies of the panel populations (F={args.warm_drift})",
        "config": {"workload": f"{args.workload}: eval I={I}, J={J}, K={K}; train {wl['train_cohorts']} cohorts "
                               f"I={wl['train_I']}, J={wl['train_J']}, K=4..10, alpha in [0.1, 1]",
                   "I": I, "J": J, "K": K},
        "ll_star": ll_star,
        "training": {"rows": int(X.shape[0]), "cohorts": info, "data_seconds": t_data, "train_seconds": t_train,
                     "epochs": args.train_epochs, "loss_first": float(hist[0]), "loss_last": float(hist[-1])},
        "forms": res,
    }
    print(json.dumps(out), flush=True)


def skfr_block(oa, ctx, wl, args):
    """SKFR AIM pre-selection (src/driver.jl:56-101) on the same resident BED:
    k = K clusters, sparsity = J/10 (the reference's documented example keeps
    10,000 of 54,051 SNPs), k-means++ start, then up to --skfr-iters
    iterations.  Kernel times come from the context's HIP-event profile; the
    centers pass (one read of the whole BED per iteration) is HBM-bound, the
    table-sum pass over the selected SNPs is FP64 vector-ALU bound.  The CPU
    figure is the single-threaded oracle on a bounded sample of the same
    cohort."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    I, J, K = wl["I"], wl["J"], wl["K"]
    s = max(1, J // 10)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    start = ctx.skfr_init(K, args.seed, 0)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    cl, _, aims, _, obj, iters = ctx.skfr_run(K, start, s, max_iter=args.skfr_iters, centers=False)
    t_run = time.perf_counter() - t0
    rep = ctx.profile_report()
    ctx.profile(False)
    bed_bytes = ((I + 3) // 4) * J
    n_c, ms_c = rep["skfr_centers"]
    n_t, ms_t = rep["skfr_tabsum"]
    per_c = ms_c / n_c / 1e3
    per_t = ms_t / n_t / 1e3
    blk = {
        "config": f"I={I}, J={J}, k={K}, sparsity={s}, global ranking, k-means++ seed {args.seed}",
        "iterations": iters, "init_seconds": t_init, "run_seconds": t_run,
        "ms_per_iteration": t_run / iters * 1e3,
        "genotype_iters_per_s": I * J * iters / t_run,
        "kernels_ms": {k: round(t / n, 4) for k, (n, t) in rep.items() if k.startswith("skfr")},
        "roofline_centers": {"bound": "hbm", "achieved": bed_bytes / per_c / 1e9, "peak": 8000.0, "unit": "GB/s",
                             "frac": bed_bytes / per_c / 8e12,
                             "note": f"algorithmic bytes = the BED, ceil(I/4)*J = {bed_bytes} per launch"},
        "tabsum": {"lookups_per_s": I * s * K / per_t, "avg_ms": per_t * 1e3,
                   "note": "one FP64 table lookup + add per (individual, selected SNP, cluster)"},
        "objective": obj,
    }
    if not args.no_cpu_baseline:
        import pyoracle as O

        from importlib import import_module

        sim = import_module("openadmixture_jl_amd.sim")
        Is, Js = min(I, 4000), min(J, 4000)
        Qt, Pt = sim.draw_truth(Is, Js, K, wl["alpha"], wl["p_dist"], wl["p_lo"], wl["p_hi"], seed=args.seed)
        pitch = (Is + 3) // 4
        bed = O.gen_bed(Qt, Pt, args.seed, wl["miss_rate"], pitch)
        st, _ = O.skfr_init(bed, Is, Js, K, args.seed, 1)
        t0 = time.perf_counter()
        r = O.skfr_run(bed, Is, Js, K, st, max(1, Js // 10), "global", 3, 1)
        el = time.perf_counter() - t0
        blk["cpu_baseline"] = {"value": Is * Js * r[5] / el, "unit": "genotype*iter/s", "cores": 1, "kind": "port",
                               "sample": f"SKFR oracle (C), I={Is}, J={Js}, k={K}, sparsity={Js // 10}, "
                                         f"{r[5]} iterations, {el:.2f} s"}
        blk["gpu_over_cpu"] = blk["genotype_iters_per_s"] / blk["cpu_baseline"]["value"]
    return blk


def calibrate_copy(nbytes=1 << 30):
    """A plain device-to-device copy of known size (HIP runtime copy kernel), used
    to scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)


def roofline_of(rep, form, K, I, J, tuning, args):
    """Dominant pass kernel of a profiled run and the FLOP rate of the algebra
    it executes against the peak of the unit that runs it (FP64 / non-packed
    FP32 vector ALU, or the BF16 matrix cores for the BF16 form's matrix-core
    update passes)."""
    dom = max((k for k in rep if k in EXEC_FLOPS), key=lambda k: rep[k][1])
    n_launch, ms = rep[dom]
    per_launch_s = ms / n_launch / 1e3
    flops = EXEC_FLOPS[dom](K)
    # the plain form's genotype passes run in FP64 for FP64 and FP32 contexts alike
    vunit = "f64" if form == "plain" else "f32"
    peak = PEAK_VALU_TFLOPS[vunit]
    bound = "valu"
    if form == "bf16" and dom in ("qside_update_bf16", "pside_update_bf16"):
        side = "q" if dom.startswith("q") else "p"
        flag = (tuning or {}).get("mfma_" + side, -1)
        if flag == -1:
            flag = args.qside_mfma if side == "q" else args.pside_mfma
        if flag == 1 or (flag in (None, -1) and K >= 7):  # the library's default rule
            flops, peak, bound, vunit = mfma_pass_flops(dom, K), PEAK_MFMA_BF16_TFLOPS, "mfma", "bf16 matrix"
    achieved = flops * I * J / per_launch_s / 1e12
    frac = achieved / peak
    if not frac <= 1.0:
        sys.exit(f"roofline fraction {frac:.3f} > 1 for {dom}: the executed-FLOP count is wrong")
    return dict(kernel=dom, flops=flops, peak=peak, bound=bound, vunit=vunit, achieved=achieved, frac=frac,
                per_launch_s=per_launch_s)


def bf16_form_block(oa, ctx, wl, args):
    """The BF16 form (update passes, log-likelihood and EM sums on the matrix
    cores; FP64 log-likelihood reduction, QP and convergence test) on the same
    cohort and context: EM(5) priming and warmup + steps iterations from the
    same seeded start as the plain line, timed the same way, with the roofline
    of its dominant kernel against the BF16 matrix peak."""
    I, J, K = wl["I"], wl["J"], wl["K"]
    ctx.set_precision("bf16")
    try:
        d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(args.seed + 1))
        oa.init_em(d, None, 5, d_cu=ctx, log=None)
        ctx.set_uv(d.U, d.V)
        it = 0
        for _ in range(args.warmup):
            it += 1
            ctx.qn_iteration(it)
        ctx.profile(True)
        ctx.profile_reset()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            it += 1
            ctx.qn_iteration(it)
        ctx.synchronize()
        el = time.perf_counter() - t0
        rep = ctx.profile_report()
        ctx.profile(False)
    finally:
        ctx.set_precision("f64")
    roof = roofline_of(rep, "bf16", K, I, J, None, args)
    dom = roof["kernel"]
    return {
        "value": I * J * args.steps / el, "unit": "genotype*iter/s", "ms_per_step": el / args.steps * 1e3,
        "steps": args.steps, "dtype": "bf16 passes, f64 LL/QP",
        "roofline": {"bound": roof["bound"], "kernel": dom, "achieved": roof["achieved"], "peak": roof["peak"],
                     "unit": "TFLOP/s", "frac": roof["frac"],
                     "traffic": measured_traffic(dom + ("_mfma" if roof["bound"] == "mfma" else ""), args.workload,
                                                 "bf16"),
                     "flops_per_genotype": roof["flops"], "avg_ms": roof["per_launch_s"] * 1e3,
                     "note": f"{roof['flops']} executed {roof['vunit']} flops/genotype of the dominant kernel x "
                             f"{I * J} genotypes per launch / its mean HIP-event time"},
        "kernels": {k: {"launches": n, "avg_ms": round(t / n, 4), "share": round(t / (el * 1e3), 4)}
                    for k, (n, t) in sorted(rep.items(), key=lambda kv: -kv[1][1])[:8]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--form", default="plain", choices=["plain", "f32", "bf16"],
                    help="update-pass precision of a FP64 run: plain FP64, FP32 vector ALU, or BF16 matrix cores")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-full-snps", action="store_true",
                    help="time the CPU oracle on all J SNPs (two full iterations; minutes at c3) instead of a sample")
    ap.add_argument("--sweep-workers", type=int, default=1, help="K values fitted at once in the c4 sweep")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run a 1 GiB device copy (counter scaling)")
    ap.add_argument("--target-waves", type=int, default=0, help="waves per genotype pass (0: library default)")
    ap.add_argument("--no-ttc", action="store_true", help="skip the time-to-converge runs")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the BF16 form's throughput block of the plain line")
    ap.add_argument("--graphs", action="store_true", help="replay steady-state iterations from hipGraphs")
    ap.add_argument("--qside-mfma", type=int, default=None, choices=[-1, 0, 1],
                    help="BF16 form's update_q pass: 1 all products on the matrix cores, 0 Hessian only")
    ap.add_argument("--pside-mfma", type=int, default=None, choices=[-1, 0, 1],
                    help="BF16 form's update_p pass: 1 all products on the matrix cores, 0 Hessian only")
    ap.add_argument("--autotune", action="store_true",
                    help="pick waves per pass and the Hessian split with the autotuner before timing")
    ap.add_argument("--stream-snps", type=int, default=0,
                    help="keep the BED in pinned host memory and stream it in chunks of this many SNPs")
    ap.add_argument("--stream-from-file", action="store_true",
                    help="with --stream-snps: write the cohort as a PLINK .bed and stream from the file")
    ap.add_argument("--resident-snps", type=int, default=0,
                    help="with --stream-snps: keep the first this many SNPs in HBM")
    ap.add_argument("--ttc-forms", default="bf16,warm_f64,warm_bf16",
                    help="forms timed to convergence beside plain FP64 (bf16, f32, warm_f64, warm_bf16)")
    ap.add_argument("--ttc-max-iter", type=int, default=200)
    ap.add_argument("--folds", type=int, default=5, help="CV folds of the K sweep (workload c4)")
    ap.add_argument("--sweep-ks", default="", help="comma-separated K values of the sweep (default 2..12)")
    ap.add_argument("--sweep-max-iter", type=int, default=300)
    ap.add_argument("--fold-start", default="full", choices=["full", "random"],
                    help="CV fold fits start from the full fit's estimates or from their own random start")
    ap.add_argument("--train-epochs", type=int, default=20, help="network training epochs (workload c5)")
    ap.add_argument("--warm-drift", type=float, default=0.05,
                    help="F of the Balding-Nichols drift between the warm start's panel populations and the "
                         "evaluation cohort's (0: the same populations)")
    ap.add_argument("--warm-p-steps", type=int, default=None,
                    help="fixed-Q P refinement iterations of the warm start (default: the library's)")
    ap.add_argument("--warm-q-steps", type=int, default=None,
                    help="fixed-P update_q! steps of an accepted warm start (default: the library's)")
    ap.add_argument("--K", type=int, default=0, help="override the workload's K (throughput runs)")
    ap.add_argument("--skfr-iters", type=int, default=10, help="SKFR iterations of the AIM pre-selection block")
    ap.add_argument("--no-skfr", action="store_true", help="skip the SKFR AIM pre-selection block")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding "
              "is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c4":
        run_sweep(args)
        return
    if args.workload == "c5":
        run_warmstart(args)
        return

    import __graft_entry__ as g

    oa = g.load_package()
    from importlib import import_module

    sim = import_module("openadmixture_jl_amd.sim")
    wl = dict(WORKLOADS[args.workload])
    if args.K:
        wl["K"] = args.K
    I, J, K = wl["I"], wl["J"], wl["K"]
    T = np.float64 if args.dtype == "f64" else np.float32

    if (args.stream_from_file or args.resident_snps) and not args.stream_snps:
        sys.exit("--stream-from-file and --resident-snps need --stream-snps")
    if args.form != "plain" and args.dtype != "f64":
        sys.exit("--form f32/bf16 runs on a FP64 context (FP64 log-likelihood, QP and convergence test)")
    ctx = oa.HIPAdmixData(I, J, K, 3, T, stream_snps=args.stream_snps)
    t0 = time.perf_counter()
    Qt_true, Pt_true = sim.cohort_on_device(ctx, wl["alpha"], wl["p_dist"], wl["p_lo"], wl["p_hi"], wl["miss_rate"],
                                            seed=args.seed)
    t_gen = time.perf_counter() - t0
    stream_note = ""
    if args.stream_snps:
        stream_note = f", streamed in {args.stream_snps}-SNP chunks from pinned host memory"
    if args.stream_from_file:
        import atexit
        import shutil
        import tempfile

        tmpd = tempfile.mkdtemp(prefix="admx_bench_")
        atexit.register(shutil.rmtree, tmpd, True)
        path = oa.write_plink(os.path.join(tmpd, "cohort"), ctx.download_bed(), I, J)
        ctx.close()
        ctx = oa.HIPAdmixData(I, J, K, 3, T, stream_snps=args.stream_snps, bed_file=path)
        stream_note = f", streamed in {args.stream_snps}-SNP chunks from a .bed file (page cache)"
    if args.resident_snps:
        ctx.set_resident_snps(args.resident_snps)
        stream_note += f", SNPs 0..{min(args.resident_snps, J) - 1} resident in HBM"
    if args.target_waves:
        ctx.set_tuning(oa._native.ADMX_TUNE_TARGET_WAVES, args.target_waves)
    if args.graphs:
        ctx.set_graphs(True)
    if args.form != "plain":
        ctx.set_precision(args.form)
    if args.qside_mfma is not None:
        ctx.set_tuning(oa._native.ADMX_TUNE_QSIDE_MFMA, args.qside_mfma)
    if args.pside_mfma is not None:
        ctx.set_tuning(oa._native.ADMX_TUNE_PSIDE_MFMA, args.pside_mfma)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(args.seed + 1), T=T)
    oa.init_em(d, None, 5, d_cu=ctx, log=None)
    ctx.set_uv(d.U, d.V)
    tuning = None
    if args.autotune:
        tuning = oa.autotune(ctx)  # on the primed state; writes x_next only
    it = 0
    for _ in range(args.warmup):
        it += 1
        ctx.qn_iteration(it)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    lls = []
    for _ in range(args.steps):
        it += 1
        lls.append(ctx.qn_iteration(it)["new"])
    ctx.synchronize()
    el = time.perf_counter() - t0
    rep = ctx.profile_report()
    ctx.profile(False)

    value = I * J * args.steps / el
    roof = roofline_of(rep, args.form, K, I, J, tuning, args)
    dom, flops, peak, bound, vunit, achieved, frac, per_launch_s = (
        roof[k] for k in ("kernel", "flops", "peak", "bound", "vunit", "achieved", "frac", "per_launch_s"))
    kernels = {k: {"launches": n, "avg_ms": round(t / n, 4), "share": round(t / (el * 1e3), 4)}
               for k, (n, t) in sorted(rep.items(), key=lambda kv: -kv[1][1])}
    out = {
        "metric": "genotype*iterations/sec at K=8 (admixture_qn! ZAL iteration)",
        "value": value,
        "unit": "genotype*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if args.dtype == "f64" else "f64 passes, f32 state and QP") if args.form == "plain"
        else f"{args.form} passes, f64 LL/QP",
        "data": "synthetic cohort generated on device (Dirichlet Q, Beta/Uniform P, Binomial(2, q.p) genotypes)",
        "config": {"workload": f"{args.workload}: I={I}, J={J}, K={K}, Q=3, ZAL, {args.form} {args.dtype}{stream_note}",
                   "I": I, "J": J, "K": K, "genotypes": I * J, "bed_bytes": ((I + 3) // 4) * J},
        "roofline": {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": frac,
                     "traffic": None if args.stream_snps else
                     measured_traffic(dom + ("_mfma" if bound == "mfma" else ""), args.workload,
                                      args.dtype if args.form == "plain" else args.form),
                     "flops_per_genotype": flops, "avg_ms": per_launch_s * 1e3,
                     "note": f"{flops} executed {vunit} flops/genotype (FMA = 2; "
                             f"{'rank-1 upper-triangle form' if dom == 'qside_update' else 'kernel algebra'}) x "
                             f"{I * J} genotypes per launch / the kernel's mean HIP-event time; "
                             + (f"{vunit} vector-ALU peak (MI355X FP64 78.6 TF; FP32 78.6 TF without packed FP32)"
                                if bound == "valu" else "MI355X dense BF16 matrix peak 2.5 PF; the weights "
                                "(about 20 vector instructions per genotype) run beside the matrix cores")
                             + (f"; the reference's full two-term count is {REF_FLOPS[dom](K)} flops/genotype"
                                if dom in REF_FLOPS else "")},
        "kernels": kernels,
        "ll_last": lls[-1],
        "gen_seconds": t_gen,
        "tuning": tuning,
    }
    if args.form == "plain" and args.dtype == "f64" and not args.stream_snps and not args.no_forms:
        out["bf16_form"] = bf16_form_block(oa, ctx, wl, args)
    if not args.no_ttc and args.dtype == "f64":
        tol = 1e-4
        hist, its = converge_run(oa, ctx, wl, T, args.seed, "f64", args.ttc_max_iter)
        ll_star = hist[-1][1]
        ttc = {"criterion": f"|LL - LL*| <= {tol} |LL*|; LL* = final LL of the plain FP64 run (rtol 1e-7, "
                            f"at most {args.ttc_max_iter} iterations)",
               "ll_star": ll_star, "plain_iterations_run": its,
               "plain_converged": its < args.ttc_max_iter}
        if I * J <= 5e7 and not args.no_cpu_baseline:
            # small enough to run the CPU oracle to convergence from the same start:
            # LL* becomes the oracle's final LL, and the CPU time to it is measured
            oc = oracle_converge(ctx, wl, args.seed, args.ttc_max_iter, tol)
            ttc["ll_star_gpu_plain"] = ll_star
            ll_star = oc["ll_final"]
            ttc["ll_star"] = ll_star
            ttc["criterion"] = (f"|LL - LL*| <= {tol} |LL*|; LL* = final LL of the FP64 oracle (CPU restatement) "
                                f"from the same start (rtol 1e-7)")
            ttc["cpu_measured"] = oc
        hit = first_within(hist, ll_star, tol)
        ttc["plain"] = {"seconds": hit[1], "qn_iterations": hit[0], "ll": hit[2],
                        "ms_per_iteration": (hist[-1][0] - hist[0][0]) / max(its, 1) * 1e3,
                        "priming_seconds": hist[0][0]}
        forms = [f for f in args.ttc_forms.split(",") if f]
        net = panel = None
        if any(f.startswith("warm") for f in forms):
            # the network is trained once on cohorts simulated from the populations
            # of this workload (subsets of K_c = 4..K of them, I=5,000 over the
            # first 20,000 SNPs), each solved by the plain method
            tw = time.perf_counter()
            Jp = min(20_000, J)
            # panel populations: Balding-Nichols drifts (F = --warm-drift) of the cohort's
            # populations, so the network never sees the evaluation cohort's own
            # allele frequencies (0: the cohort's populations themselves)
            P_panel = (np.asfortranarray(Pt_true) if args.warm_drift <= 0 else
                       oa.drifted_populations(Pt_true, args.warm_drift, args.seed + 17))
            panel = oa.Panel(P_panel[:, :Jp].mean(0), seed=args.seed)
            X, Tt, _ = oa.training_set(P_panel, panel, 10, I=5_000, J=Jp, Ks=range(min(4, K), K + 1),
                                       alpha_range=(0.1, 1.0), miss_rate=wl["miss_rate"], seed=args.seed)
            t_data = time.perf_counter() - tw
            net = oa.MLP(64, 256, K, seed=args.seed)
            tw = time.perf_counter()
            lh = net.train(X, Tt, epochs=20, batch=512, lr=2e-3, seed=args.seed)
            ttc["warm_start_training"] = {
                "panel": (f"Balding-Nichols drifts (F = {args.warm_drift}) of the cohort's populations"
                          if args.warm_drift > 0 else "the cohort's own populations"),
                "rows": int(X.shape[0]), "cohorts": 10, "simulate_and_solve_seconds": t_data,
                "train_seconds": time.perf_counter() - tw, "loss_first": float(lh[0]), "loss_last": float(lh[-1]),
                "note": "one-time cost, not part of any time to converge"}
        for form in forms:
            if form.startswith("warm"):
                prec = form.split("_", 1)[1] if "_" in form else "f64"
                ctx.set_precision(prec)
                r = oa.warm_started_fit(ctx, net, panel, seed=args.seed + 1, max_iter=args.ttc_max_iter, **_warm_kw(args),
                                        stop_at=(ll_star, tol))
                h2, its2 = r["history"], r["iterations"]
            else:
                h2, its2 = converge_run(oa, ctx, wl, T, args.seed, form, args.ttc_max_iter, stop_at=(ll_star, tol))
            hit2 = first_within(h2, ll_star, tol)
            ttc[form] = ({"seconds": hit2[1], "qn_iterations": hit2[0], "ll": hit2[2],
                          "ms_per_iteration": (h2[-1][0] - h2[0][0]) / max(its2, 1) * 1e3,
                          "priming_seconds": h2[0][0], "vs_plain_speedup": hit[1] / hit2[1]} if hit2 else
                         {"seconds": None, "note": f"not within {tol} of LL* after {its2} iterations",
                          "ll_last": h2[-1][1]})
            if form.startswith("warm"):
                ttc[form]["chosen_start"] = r["chosen"]
                ttc[form]["judge_gain"] = r["gain"]
        if net is not None:
            net.close()
        ctx.set_precision("f64" if args.form == "plain" else args.form)
        out["time_to_converge"] = ttc
    if not args.no_skfr and args.stream_snps == 0:
        out["skfr"] = skfr_block(oa, ctx, wl, args)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(wl, args.dtype, args.seed, args.cpu_budget, full_snps=args.cpu_full_snps,
                          ctx=None if args.stream_snps else ctx)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "cpu_model")}
        out["gpu_over_cpu"] = value / cb["value"]
        if "time_to_converge" in out:
            ttc = out["time_to_converge"]
            n_it = ttc["plain"]["qn_iterations"]
            cpu_s = 5 * cb["em_step_seconds_full"] + n_it * cb["iteration_seconds_full"]
            ttc["cpu_extrapolated"] = {
                "seconds": cpu_s,
                "basis": f"oracle seconds per EM step and per iteration on {cb['cores']} host threads, scaled to "
                         f"all {J} SNPs, x (5 EM steps + {n_it} iterations of the plain GPU run)"}
            if "cpu_measured" in ttc:
                cpu_s = ttc["cpu_measured"]["seconds_to_tol"]
                ttc["cpu_extrapolated"]["note"] = "not used: the CPU oracle's time to LL* is measured (cpu_measured)"
            ttc["gpu_plain_vs_cpu"] = cpu_s / ttc["plain"]["seconds"]
            for form in args.ttc_forms.split(","):
                if form in ttc and ttc[form].get("seconds"):
                    ttc[form]["vs_cpu"] = cpu_s / ttc[form]["seconds"]
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
