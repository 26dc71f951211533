#!/usr/bin/env python3
"""DA-GTM E+M iterations/s on one MI355X (mixed precision vs FP64 GPU).

Workload (the headline configuration, --config c3): N = 1,000,000 binary fingerprint
vectors, D = 166 (50 prototypes, per-bit p ~ Beta(0.5, 2), flip 0.05,
seed 7); L = 3 latent grid K = 20^3 = 8000; M = 5^3 = 125 RBF + bias
(width = centre spacing, s = 1); ridge lambda = 1e-3.  One step = one E+M
iteration (E-pass 1, E-pass 2, reductions, M-step with the beta update) in
deterministic-annealing mode at temperature 2.0, the three-exponential
iteration that dominates an annealing run.  Inputs are resident in HBM before
the timed region; every library call is synchronous, so wall time around the
K timed steps is the device time.

Prints one JSON line (rank 0, N = 1 GPU only).

`--batch` instead measures the batched-instances workload: 64 independent
instances (N log-uniform in [1e3, 2e5]; D from {12, 32, 64, 166}, Gaussian-
mixture float data for D <= 64 and binary fingerprints for D = 166; (L, K, M)
from {(2, 400, 25), (2, 1600, 25), (3, 8000, 125)}; independent seeds and
start temperatures) annealed concurrently by dg_batch_anneal, each for a fixed
number of E+M iterations.  Its JSON line reports instance-iterations/s summed
over the batch.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "da-gtm_amd"))

import dagtm_gpu as dg  # noqa: E402
import datagen  # noqa: E402

METRIC = "DA-GTM E+M iter/s at N=1M,K=8k,D=166: mixed-prec vs FP64 GPU, same loglik"
PEAK_FP16_TFLOPS = 2500.0      # MI355X dense FP16/BF16 MFMA
PEAK_FP64_TFLOPS = 78.6        # MI355X FP64
PEAK_HBM_GBPS = 8000.0
TEMP = 2.0
LAMBDA = 1e-3

# single-instance configurations: (N, D, K, M, L, data, description)
CONFIGS = {
    "c1": (1000, 12, 400, 25, 2, "gmm",
           "N=1000 Gaussian mixture D=12 (3 components, seed 42), L=2 K=400, M+1=26, s=1, "
           "lambda=1e-3, DA temp 2.0; one step = one E+M iteration"),
    "c2": (100_000, 166, 8000, 125, 3, "fp50",
           "N=100,000 binary fingerprints D=166 (50 prototypes, flip 0.05, seed 7), L=3 K=8000, "
           "M+1=126, s=1, lambda=1e-3, DA temp 2.0; one step = one E+M iteration"),
    "c3": (1_000_000, 166, 8000, 125, 3, "fp50",
           "N=1,000,000 binary fingerprints D=166 (50 prototypes, flip 0.05, seed 7), L=3 K=8000, "
           "M+1=126, s=1, lambda=1e-3, DA temp 2.0; one step = one E+M iteration"),
    "c5": (2_000_000, 166, 64000, 1000, 3, "fp200",
           "N=2,000,000 binary fingerprints D=166 (200 prototypes, flip 0.05, seed 7), L=3 "
           "K=64000, M+1=1001, s=1, lambda=1e-3, DA temp 2.0; one step = one E+M iteration"),
}
# CPU baseline samples per configuration: (points at all threads, points at
# one thread); None = the whole workload, 0 = not measured.  A sample smaller
# than the workload is timed at n and n / 2 points and the per-iteration time
# a + b n extrapolated to the workload's N (the M-step's A = FI^T G FI and
# solve do not grow with N; the rest does).
CPU_SAMPLES = {"c1": (None, None), "c2": (None, 4000), "c3": (16000, 2000), "c5": (1000, 0)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary; None when there is none or when the
    kernel source changed since it was measured (sha256 recorded beside it)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{kernel}_bench_default.md")
    src = os.path.join(ROOT, "da-gtm_amd", "csrc", "dg_mixed.hip")
    try:
        rec = {}
        for line in open(path):
            if ":" in line and line.split(":", 1)[0] in ("traffic_bytes_per_launch", "source_sha256"):
                k, v = line.split(":", 1)
                rec[k] = v.split()[0]
        sha = hashlib.sha256(open(src, "rb").read()).hexdigest()
        if rec.get("source_sha256") != sha:
            return None
        return float(rec["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def build_problem(N, D, K, M, L, data, seed=7):
    if data == "gmm":
        T = datagen.gmm(N, D, 42)
    else:
        T = datagen.fingerprints(N, D, seed, nproto=200 if data == "fp200" else 50)
        T = T.astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, X, FI, Y0, b0


def run_steps(eng, steps, visited=None):
    """steps E+M iterations; returns (seconds, per-step llh, summed phase ms).
    With a list `visited` (mixed engines), appends each step's fraction of
    (centre group, block) pairs that E-pass 2 computed."""
    llh = []
    phase = np.zeros(4)
    probe = 0.0                    # host time of the visited-pairs read-back, not part of a step
    t0 = time.perf_counter()
    for _ in range(steps):
        l, _, _ = eng.estep(True, TEMP)
        ph = eng.timing()
        if visited is not None:
            # the E-step has returned (the stream is idle): this read-back only
            # adds host time, which is taken out of the step time
            tp = time.perf_counter()
            v, t = eng.pass2_blocks()
            visited.append(v / t)
            probe += time.perf_counter() - tp
        eng.mstep()
        ph[3] = eng.timing()[3]
        phase += ph
        llh.append(l)
    return time.perf_counter() - t0 - probe, llh, phase


def _oracle_time(O, T, FI, Y0, b0, n, threads, solver, iters=2):
    """Seconds per E+M iteration of the oracle on the first n points (the
    first iteration untimed when more than one is run)."""
    O.set_threads(threads)
    m = O.Model(np.ascontiguousarray(T[:n]), FI, Y0, b0, lam=LAMBDA, solver=solver)
    if iters > 1:
        m.estep(1, TEMP)
        m.mstep()
        iters -= 1
    t0 = time.perf_counter()
    for _ in range(iters):
        m.estep(1, TEMP)
        m.mstep()
    return (time.perf_counter() - t0) / iters


def _oracle_rate_at(O, T, FI, Y0, b0, n, threads, solver):
    """(seconds per iteration extrapolated to the whole workload, note)."""
    N = T.shape[0]
    if n >= N:
        t = _oracle_time(O, T, FI, Y0, b0, N, threads, solver)
        return t, f"the whole workload, {t:.3f} s per iteration"
    t_hi = _oracle_time(O, T, FI, Y0, b0, n, threads, solver)
    t_lo = _oracle_time(O, T, FI, Y0, b0, n // 2, threads, solver)
    b = max(t_hi - t_lo, 0.0) / (n - n // 2)
    a = max(t_hi - b * n, 0.0)
    return a + b * N, (f"the first {n} and {n // 2} points ({t_hi:.3f} and {t_lo:.3f} s per "
                       f"iteration), a + b N fitted and extrapolated to N={N}")


def cpu_baseline(cfg, T, FI, Y0, b0):
    """The oracle (FP64, the reference's step order, materialised D and R) on
    this box's host cores: all the threads this process may use (16 on a
    one-GPU box) and one thread, each on a bounded sample of the workload.
    Also compares one E+M iteration of both GPU paths with the oracle on the
    all-thread sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    N = T.shape[0]
    ns, ns1 = CPU_SAMPLES[cfg]
    ns = N if ns is None else min(ns, N)
    # the oracle's pseudo-inverse (Jacobi eigen-decomposition) of a 1001 x 1001
    # A takes minutes; at M + 1 > 126 the timed solve is its Cholesky
    solver = O.SOLVE_PINV if FI.shape[1] <= 126 else O.SOLVE_CHOL
    threads = min(len(os.sched_getaffinity(0)), 16)
    dt, note = _oracle_rate_at(O, T, FI, Y0, b0, ns, threads, solver)
    # GPU paths on the same sample, one iteration from the same state
    m1 = O.Model(np.ascontiguousarray(T[:ns]), FI, Y0, b0, lam=LAMBDA, solver=O.SOLVE_CHOL)
    llh_o, _, _ = m1.estep(1, TEMP)
    m1.mstep()
    beta_o = m1.beta
    del m1
    vs = {}
    for name, prec in (("fp64", dg.FP64), ("mixed", dg.MIXED)):
        with dg.Engine(precision=prec, lam=LAMBDA) as e:
            e.upload(T[:ns], FI, Y0, b0)
            llh, _, _ = e.estep(True, TEMP)
            beta = e.mstep()
        vs[name] = {"llh_rel_err": abs(llh - llh_o) / abs(llh_o),
                    "beta_rel_err": abs(beta - beta_o) / beta_o}
    single = None
    if ns1 != 0:
        ns1 = N if ns1 is None else min(ns1, N)
        dt1, note1 = _oracle_rate_at(O, T, FI, Y0, b0, ns1, 1, solver)
        single = {"value": 1.0 / dt1, "unit": "iter/s", "sample": note1}
    return {"value": 1.0 / dt, "unit": "iter/s", "cores": threads, "kind": "port",
            "sample": f"oracle/liboracle.so (FP64, reference step order, materialised D and R, "
                      f"{'pseudo-inverse' if solver == O.SOLVE_PINV else 'Cholesky'} solve) "
                      f"with {threads} threads on {note}; K={FI.shape[0]}, D={T.shape[1]}",
            "nproc": os.cpu_count(), "threads_allowed": len(os.sched_getaffinity(0)),
            "single_thread": single,
            "gpu_vs_oracle_on_sample": {"points": ns, **vs}}


def device_peaks():
    with dg.Engine(precision=dg.FP64) as e:
        return e.device_peaks()


def batch_bench(args):
    """Instance-iterations/s of the batched workload: the annealing phase of
    dg_batch_anneal (every instance's data already resident), timed by the
    library, after one untimed batch of `warmup` iterations."""
    print("[bench] generating 64 instances", file=sys.stderr, flush=True)
    probs = datagen.batch_instances()
    n = len(probs)
    prec = dg.MIXED if args.batch_precision == "mixed" else dg.FP64
    tuned = 0
    if args.tune and prec != dg.FP64:
        import tempfile
        table = os.path.join(tempfile.mkdtemp(), "dg_tune.txt")
        seen = set()
        for T, FI, Y0, b0, _ in probs:
            key = (T.shape[1], FI.shape[0], int(np.ceil(np.log2(T.shape[0]))))
            if key in seen:
                continue
            seen.add(key)
            with dg.Engine(precision=prec, lam=LAMBDA) as e:
                e.upload(T, FI, Y0, b0)
                e.tune(table)
            tuned += 1
        print(f"[bench] tuned {tuned} shapes", file=sys.stderr, flush=True)
        os.environ["DG_TUNE_FILE"] = table
    for iters in (args.warmup, args.steps):
        print(f"[bench] batch run, {iters} iterations per instance", file=sys.stderr, flush=True)
        out, ph = dg.batch_anneal(probs, precision=prec, lam=LAMBDA, workers=args.workers,
                                  alpha=0.99, eps=0.0, maxloop=iters)
        bad = [i for i, o in enumerate(out) if o[0] != dg.DG_OK or o[1].nloop != iters]
        if bad:
            print(f"[bench] instances failed: {bad}", file=sys.stderr)
            sys.exit(3)
    dt = ph[1] * 1e-3
    knd = sum(p[1].shape[0] * p[0].shape[0] * p[0].shape[1] for p in probs)
    f_alg = 4.0 * knd * args.steps / dt / 1e12         # algorithmic TFLOP/s of the E-steps
    pk_mfma = device_peaks()[0] if prec != dg.FP64 else None
    if args.workers <= 0 and prec != dg.FP64:
        mode = ("batched launches per step, " + ("replayed as hipGraphs" if args.workers == 0
                                                  else "launched directly") +
                "; E-passes as persistent launches over a device work queue")
        eff_workers = 1
    else:
        eff_workers = min(4 if args.workers <= 0 else min(args.workers, 4), n)
        mode = f"{eff_workers} host worker threads, one context and stream each"
    tbytes = sum(p[0].shape[0] * p[0].shape[1] * (1 if p[0].shape[1] == 166 else 4) for p in probs)
    it_s = n * args.steps / dt
    hbm = tbytes * args.steps / dt / 1e9
    print(json.dumps({
        "metric": "batched DA-GTM instance E+M iter/s (64 instances of uneven N, D, K)",
        "value": round(it_s, 2), "unit": "instance-iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_batch_step": round(1e3 * dt / args.steps, 3),
        "upload_ms": round(ph[0], 1),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp16x2-mfma+fp32 (E-step), fp64 (M-step)" if prec == dg.MIXED else "fp64",
        "data": "synthetic",
        "config": {"workload": "64 instances: N log-uniform [1e3, 2e5], D in {12,32,64,166}, "
                               "(L,K,M) in {(2,400,25),(2,1600,25),(3,8000,125)}, seed 2024",
                   "total_N": int(sum(p[0].shape[0] for p in probs)), "workers": args.workers,
                   "workers_effective": eff_workers, "tuned_shapes": tuned,
                   "parallelism": "single GPU, " + mode},
        "kd_updates_per_s": knd * args.steps / dt,
        "mfma": {"algorithmic_tflops": round(f_alg, 2), "peak_tflops": PEAK_FP16_TFLOPS,
                 "frac": round(f_alg / PEAK_FP16_TFLOPS, 4),
                 "measured_peak_tflops": None if pk_mfma is None else round(pk_mfma, 1),
                 "frac_of_measured": None if pk_mfma is None else round(f_alg / pk_mfma, 4),
                 "note": "F_alg = 4 K N D flop per instance-iteration"},
        "data_bytes_per_batch_step": tbytes,
        "hbm": {"achieved_GBps_data_once": round(hbm, 2), "peak_GBps": 8000.0,
                "frac": round(hbm / 8000.0, 5),
                "note": "the workload is MFMA/VALU-bound (K*N*D work per data byte >= 400), "
                        "so HBM is not its binding roof"},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", choices=sorted(CONFIGS), default="c3",
                    help="c3 (default): the headline workload; c1, c2, c5: the other "
                         "single-instance configurations (c4 is --batch)")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["mixed", "mixed_r1"], default="mixed",
                    help="mixed: R as an FP16 hi/lo pair; mixed_r1: R as one FP16 term")
    ap.add_argument("--batch", action="store_true",
                    help="measure the 64-instance batched workload instead of the headline")
    ap.add_argument("--batch-precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--workers", type=int, default=0,
                    help="with --batch: 0 = batched launches as hipGraphs, -1 = batched launches "
                         "without graphs, 1..4 = host worker threads")
    ap.add_argument("--tune", action="store_true",
                    help="with --batch: tune the pass-2 chunk count of every distinct shape first "
                         "(dg_tune) and run the batch with that table")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = (1 << 30) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; sharding N over several GPUs (one "
              "all-reduce of G and R T per iteration) is designed but not built, so --gpus "
              "must be 1", file=sys.stderr)
        sys.exit(2)
    if args.batch:
        batch_bench(args)
        return

    cfg = args.config
    N, D, K, M, L, data, desc = CONFIGS[cfg]
    if args.N:
        N = args.N
    fp64_steps = args.fp64_steps or (1 if cfg == "c5" else 3)
    print(f"[bench] {cfg}: generating N={N} D={D} K={K} M+1={M + 1}", file=sys.stderr, flush=True)
    T, X, FI, Y0, b0 = build_problem(N, D, K, M, L, data)

    # ---- mixed precision (the measured path)
    prec16 = dg.MIXED_R1 if args.precision == "mixed_r1" else dg.MIXED
    e16 = dg.Engine(precision=prec16, lam=LAMBDA)
    t_up = time.perf_counter()
    e16.upload(T, FI, Y0, b0)            # host buffers cross PCIe once, here
    upload_s = time.perf_counter() - t_up
    e16.set_timing(True)
    print("[bench] mixed warmup", file=sys.stderr, flush=True)
    # the state after the first iteration is compared with the FP64 path's
    _, llh16_w, _ = run_steps(e16, 1)
    st16 = e16.download()
    _, llh_more, _ = run_steps(e16, max(0, args.warmup - 1))
    llh16_w += llh_more
    print("[bench] mixed timed", file=sys.stderr, flush=True)
    visited = []
    sec16, llh16_t, ph16 = run_steps(e16, args.steps, visited)
    llh16 = llh16_w + llh16_t
    it16 = args.steps / sec16
    e16.close()

    # the same upload from rows padded as a gsl_matrix pads them (64 bytes)
    upload_padded_s = None
    if cfg in ("c3", "c5"):
        ldT = (D * 8 + 63) // 64 * 8
        Tp = np.zeros((N, ldT))
        Tp[:, :D] = T
        with dg.Engine(precision=prec16, lam=LAMBDA) as ep:
            t_up = time.perf_counter()
            ep.upload_strided(Tp, D, FI, Y0, b0)
            upload_padded_s = time.perf_counter() - t_up
        del Tp

    # ---- FP64 GPU path from the same initial state
    e64 = dg.Engine(precision=dg.FP64, lam=LAMBDA)
    e64.upload(T, FI, Y0, b0)
    e64.set_timing(True)
    print("[bench] fp64", file=sys.stderr, flush=True)
    _, llh64_w, _ = run_steps(e64, 1)
    st64 = e64.download()
    sec64, llh64_t, ph64 = run_steps(e64, fp64_steps)
    llh64 = llh64_w + llh64_t
    it64 = fp64_steps / sec64
    e64.close()

    n = min(len(llh16), len(llh64))
    llh_err_steps = [abs(a - b) / abs(b) for a, b in zip(llh16[:n], llh64[:n])]
    llh_err = max(llh_err_steps)
    # W, Y (max-norm relative) and beta after one iteration from the same state
    state_err = {"W": float(np.abs(st16[0] - st64[0]).max() / np.abs(st64[0]).max()),
                 "Y": float(np.abs(st16[1] - st64[1]).max() / np.abs(st64[1]).max()),
                 "beta": abs(st16[2] - st64[2]) / st64[2]}

    # ---- roofline of the dominant kernel (per launch averages from HIP events)
    p1_ms, p2_ms = ph16[0] / args.steps, ph16[1] / args.steps
    flop = 2.0 * K * N * D          # algorithmic flops of each E-pass
    if p2_ms >= p1_ms:
        dom, dom_ms, name = "k_e2_mfma", p2_ms, "E-pass 2 (R T contraction)"
    else:
        dom, dom_ms, name = "k_e1_mfma", p1_ms, "E-pass 1 (cross term)"
    achieved = flop / (dom_ms * 1e-3) / 1e12
    # E-pass 2 computes only the (centre group, 32-point block) pairs where
    # some responsibility survives in FP16; the rest would add exact zeros
    vis = float(np.mean(visited)) if visited else 1.0
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP16_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP16_TFLOPS, 4),
            "traffic": measured_traffic(dom) if cfg == "c3" else None,
            "kernel": dom, "kernel_desc": name, "kernel_ms": round(dom_ms, 4),
            "algorithmic_flop_per_launch": flop,
            "pass2_visited_frac": round(vis, 4),
            "achieved_on_visited_blocks": (round(achieved * vis, 2) if dom == "k_e2_mfma"
                                           else round(achieved, 2))}

    # ---- step roofline: t_lb = max(F / MFMA peak, E / exp peak, B / HBM peak)
    # with the MFMA and exponential peaks measured on this device
    print("[bench] device peaks", file=sys.stderr, flush=True)
    pk_mfma, pk_exp, pk_f64 = device_peaks()
    bT = 2.0 if data != "gmm" else 4.0      # stored T width: FP16, or an FP16 hi/lo pair
    F_alg, E_alg = 4.0 * K * N * D, 3.0 * K * N
    B_alg = N * D * bT + 16.0 * K * D + 8.0 * N
    t_lb = {"mfma": F_alg / (pk_mfma * 1e12), "exp": E_alg / pk_exp,
            "hbm": B_alg / (PEAK_HBM_GBPS * 1e9)}
    bind = max(t_lb, key=t_lb.get)
    t_step = 1.0 / it16
    step_roof = {"t_lb_ms": round(1e3 * t_lb[bind], 4), "binding": bind,
                 "frac": round(t_lb[bind] / t_step, 4),
                 "t_lb_ms_each": {k: round(1e3 * v, 4) for k, v in t_lb.items()},
                 "measured_peaks": {"mfma_fp16_tflops": round(pk_mfma, 1),
                                    "exp_per_s": pk_exp, "fp64_fma_tflops": round(pk_f64, 2)},
                 "F_alg_flop": F_alg, "E_alg_exp": E_alg, "B_alg_bytes": B_alg}
    fp64_frac = F_alg * it64 / (PEAK_FP64_TFLOPS * 1e12)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg, T, FI, Y0, b0)

    out = {
        "metric": METRIC if cfg == "c3" else
        f"DA-GTM E+M iter/s at N={N},K={K},D={D} ({cfg}): mixed-prec vs FP64 GPU",
        "value": round(it16, 3), "unit": "iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 / it16, 3),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": ("fp16x2-mfma+fp32 (E-step), fp64 (M-step)" if prec16 == dg.MIXED else
                  "fp16x2 distances, fp16 R, fp32 accumulation (E-step), fp64 (M-step)"),
        "data": "synthetic",
        "config": {"workload": desc if args.N is None else desc + f" (N cut to {N})",
                   "name": cfg, "N": N, "D": D, "K": K, "M1": M + 1,
                   "parallelism": "single GPU"},
        "fp64_gpu_iter_per_s": round(it64, 4),
        "mixed_over_fp64": round(it16 / it64, 2),
        "fp64_path_frac_of_fp64_peak": round(fp64_frac, 4),
        "llh_rel_err_vs_fp64": llh_err,
        "llh_rel_err_vs_fp64_per_step": llh_err_steps,
        "pass2_visited_frac_per_step": [round(v, 4) for v in visited],
        "state_rel_err_vs_fp64_after_1_iter": state_err,
        "kd_updates_per_s": K * N * D * it16,
        "phase_ms_mixed": [round(x / args.steps, 4) for x in ph16],
        "upload_s": round(upload_s, 3),
        "upload_padded_rows_s": None if upload_padded_s is None else round(upload_padded_s, 3),
        "phase_ms_fp64": [round(x / fp64_steps, 4) for x in ph64],
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
