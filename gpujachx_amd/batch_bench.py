"""Throughput of many independent (H)SVDs in one persistent launch.

Workload: `--count` problems (default 4096), N drawn uniformly from
{64, 128, ..., 1024}, M from {N, 2N}, half SVDs (NPLUS = N) and half HSVDs
(NPLUS = N/2), FP64, G iid N(0,1) times column scales 10^U(-3,3); ALG 11
(block oriented, sorted, V accumulated).  The launch shape and strategy
family come from the library's autotuner (on a sample of the batch).  A run
restores the inputs, then times hj_batch_run; reports problems/s and the
p50/p99 of the per-problem device time and completion time.

    python -m gpujachx_amd.batch_bench [--count 4096] [--runs 2] [--no-tune | --cfg W,G,S,L]
"""
import argparse
import json
import os
import time

import numpy as np


def make_batch(torch, count, seed, alg):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    Ns = 64 * rng.integers(1, 17, count)
    Ms = Ns * rng.integers(1, 3, count)
    hyp = rng.uniform(size=count) < 0.5
    probs = []
    for k in range(count):
        n, m = int(Ns[k]), int(Ms[k])
        nplus = n // 2 if hyp[k] else n
        # column-major m x n (ld = m) is the row-major n x m transpose
        Gt = torch.randn(n, m, dtype=torch.float64, device="cuda", generator=gen)
        sc = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device="cuda", generator=gen) * 6.0 - 3.0)
        Gt.mul_(sc[:, None])
        V = torch.empty(n, n, dtype=torch.float64, device="cuda") if alg & 2 else None
        D = torch.empty(n, dtype=torch.float64, device="cuda")
        probs.append({"m": m, "n": n, "nplus": nplus, "G0": Gt, "G": torch.empty_like(Gt), "V": V, "D": D})
    return probs


def descs_of(g, probs, alg):
    return [g.BatchDesc(alg, p["m"], p["n"], p["nplus"], p["m"], p["n"] if p["V"] is not None else 0,
                        p["G"].data_ptr(), p["V"].data_ptr() if p["V"] is not None else None,
                        p["D"].data_ptr()) for p in probs]


def check_sample(torch, probs, k):
    """||V^T J V - J||_max and ||G V - U diag(sqrt|D|)||_max / ||G||_max on k problems."""
    worst_j, worst_r = 0.0, 0.0
    idx = np.linspace(0, len(probs) - 1, k).astype(int)
    for i in idx:
        p = probs[i]
        n, nplus = p["n"], p["nplus"]
        G0 = p["G0"].T          # m x n
        U = p["G"].T
        V = p["V"].T
        J = torch.cat([torch.ones(nplus, dtype=torch.float64, device="cuda"),
                       -torch.ones(n - nplus, dtype=torch.float64, device="cuda")])
        ej = (V.T @ (J[:, None] * V) - torch.diag(J)).abs().max().item()
        er = ((G0 @ V - U * torch.sqrt(p["D"].abs())).abs().max() / G0.abs().max()).item()
        worst_j, worst_r = max(worst_j, ej), max(worst_r, er)
    return worst_j, worst_r


FP64_PEAK_TFS = 78.6
HBM_PEAK_GBS = 8000.0


def visits_per_sweep(g, n1, strategy):
    """Block-pair visits of one level-2 sweep of order n1 under the batch's
    strategy choice (the preferred family where it has a construction, else
    BrentL, else mmstep): steps x n1/2."""
    pref = {0: "cycwor", 1: "BrentL", 2: "mmstep"}[int(strategy)]
    for fam in (pref, "BrentL", "mmstep"):
        try:
            return g.strat_generate(fam, n1).shape[0] * (n1 // 2)
        except ValueError:
            continue
    raise ValueError("no strategy of order %d" % n1)


def roofline(g, probs, sweeps, strategy, alg, seconds):
    """Algorithmic work of the batch and its rates.  Per block-pair visit (32
    columns, m rows, n_V = n rows of V when accumulated): Gram 2048 m flops,
    update 2048 (m + n_V); bytes 256 (3 m + 2 n_V) (Gram read, update read and
    write of the pair, V read and write), times the visits of every executed
    sweep."""
    flops = bytes_ = 0.0
    visits = 0
    cache = {}
    for p, sw in zip(probs, sweeps):
        n1 = p["n"] // 16
        if n1 not in cache:
            cache[n1] = visits_per_sweep(g, n1, strategy)
        v = cache[n1] * int(sw)
        nv = p["n"] if alg & 2 else 0
        visits += v
        flops += v * 2048.0 * (2.0 * p["m"] + nv)
        bytes_ += v * 256.0 * (3.0 * p["m"] + 2.0 * nv)
    tf = flops / seconds / 1e12
    gb = bytes_ / seconds / 1e9
    return {"algorithmic_flops": flops, "algorithmic_bytes": bytes_, "visits": visits,
            "tflops": tf, "frac_fp64_peak": tf / FP64_PEAK_TFS, "GBs": gb,
            "frac_hbm_peak": gb / HBM_PEAK_GBS,
            "peaks": "FP64 MFMA 78.6 TFLOP/s, HBM 8 TB/s (MI355X published)",
            "bound": "hbm",
            "note": "counter runs of the same workload (profiles/roofline_batch_4096.md): measured "
                    "HBM traffic 0.79x the algorithmic bytes (updates by an identity W are left "
                    "out) at 4.89 TB/s, FP64 MFMA busy 30 %: bound by HBM, the visits' latency "
                    "taking the rest"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--runs", type=int, default=2)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--alg", type=int, default=11)
    ap.add_argument("--sample", type=int, default=256, help="autotuner sample size")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--max-sweeps", type=int, default=30)
    ap.add_argument("--cfg", default=None,
                    help="launch shape waves,wg_per_cu,strategy,longest_first instead of tuning")
    args = ap.parse_args()
    os.environ.setdefault("HJ_QUIET", "1")
    import torch
    import gpujachx_amd as g
    g.configureGPU(0)
    t0 = time.time()
    probs = make_batch(torch, args.count, args.seed, args.alg)
    for p in probs:
        p["G"].copy_(p["G0"])
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    descs = descs_of(g, probs, args.alg)
    tune = None
    if args.no_tune or args.cfg:
        cfg = g.default_batch_cfg()
        cfg.max_sweeps = args.max_sweeps
        if args.cfg:
            cfg.waves, cfg.wg_per_cu, cfg.strategy, cfg.longest_first = \
                [int(x) for x in args.cfg.split(",")]
    else:
        t0 = time.time()
        cfg, rate = g.batch_autotune(descs, sample=args.sample, max_sweeps=args.max_sweeps)
        tune = {"seconds": time.time() - t0, "sample": args.sample, "sample_problems_per_s": rate}
    b = g.Batch(descs, cfg)
    times = []
    res = None
    for r in range(args.runs + 1):      # the first run is warmup
        for p in probs:
            p["G"].copy_(p["G0"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        if r:
            times.append(time.perf_counter() - t0)
        res = b.results()
    b.close()
    lat = res["end_us"] - res["start_us"]
    ej, er = check_sample(torch, probs, 32) if args.alg & 2 else (None, None)
    best = min(times)
    out = {
        "metric": "batched (H)SVD throughput, one persistent launch",
        "problems": args.count, "problems_per_s": args.count / best, "seconds": best,
        "runs": times, "alg": args.alg,
        "latency_us": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "completion_us": {"p50": float(np.percentile(res["end_us"], 50)),
                          "p99": float(np.percentile(res["end_us"], 99))},
        "sweeps": {"mean": float(res["sweeps"].mean()), "max": int(res["sweeps"].max()),
                   "at_cap": int((res["sweeps"] >= cfg.max_sweeps).sum())},
        "cfg": cfg.as_dict(), "autotune": tune,
        "roofline": roofline(g, probs, res["sweeps"], cfg.strategy, args.alg, best),
        "check_32": {"VtJV_minus_J_max": ej, "residual_rel_max": er},
        "workload": "N uniform in {64..1024} step 64, M in {N, 2N}, half SVD (NPLUS=N), half HSVD "
                    "(NPLUS=N/2), G iid N(0,1) x 10^U(-3,3) column scales, FP64",
        "setup_seconds": gen_s,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
