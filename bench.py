#!/usr/bin/env python3
"""Headline benchmark: seconds per SVD of an M=N=8192 FP64 matrix on one MI355X.

Workload: G = U diag(s) V^T with Haar U, V (QR of a seeded N(0,1) matrix,
column signs from diag R) and s_i = 10^(-12 u_i), u_i ~ U(0,1); strategies
cycwor/cycwor; ALG 9 (block oriented, sorted, sigma only).  A step is one
complete SVD (all sweeps until an unskipped sweep has no big rotation, then
the column normalisation) of a fresh copy of G already resident in HBM.

`value` is the fast form (learned pair-skip with the committed model); the
plain FP64 form, bit-identical to the oracle, is timed beside it on the same
matrix and reported as `plain`, with the ratio.

    python bench.py [--steps K] [--warmup W] [--n N] [--alg A] [--form F] [--no-cpu]

Prints one JSON line.  One GPU only: --gpus above 1 exits with a message.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s per HSVD at N=8192 (learned+split vs plain FP64, MI355X), rel sigma err<=1e-12"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def make_input(torch, n, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    dev = torch.device("cuda")

    def haar():
        a = torch.randn(n, n, dtype=torch.float64, device=dev, generator=gen)
        q, r = torch.linalg.qr(a)
        return q * torch.sign(torch.diagonal(r))

    U = haar()
    V = haar()
    u = torch.rand(n, dtype=torch.float64, device=dev, generator=gen)
    s = torch.pow(10.0, -12.0 * u)
    G = (U * s) @ V.T
    # column-major G (ld = n) is the row-major transpose
    return G.T.contiguous(), s


def cpu_baseline(G_host, n, alg, sweeps, steps_per_sweep, budget_s=15.0):
    """The oracle (C restatement, OpenMP over the pairs of a step) timed on a
    bounded sample: the first S level-1 steps of the same SVD; the per-step
    time is scaled to the GPU run's step count."""
    import oracle as o
    import gpujachx_amd as g
    s0, s1 = g.current_strats()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    t0 = time.time()
    o.hypJacL2(alg, G_host, n, s0, s1, nthreads=threads, max_steps=1)
    t1 = time.time() - t0
    S = max(1, min(steps_per_sweep, int(budget_s / max(t1, 1e-3))))
    t0 = time.time()
    o.hypJacL2(alg, G_host, n, s0, s1, nthreads=threads, max_steps=S)
    tS = time.time() - t0
    per_step = tS / S
    total_steps = sweeps * steps_per_sweep
    return {
        "value": per_step * total_steps, "unit": "s", "cores": threads, "kind": "port",
        "sample": "oracle (C, FP64, OpenMP over the pairs of a step) on the first %d of the %d "
                  "level-1 steps of the same %dx%d SVD (%.2f s), scaled by %d sweeps x %d steps"
                  % (S, total_steps, n, n, tS, sweeps, steps_per_sweep),
    }


def dgesvj_sample(n_small=1024, seed=3):
    """LAPACK dgesvj on the same construction at a smaller order, for scale."""
    from oracle.lapack import dgesvj
    rng = np.random.default_rng(seed)

    def haar(k):
        q, r = np.linalg.qr(rng.standard_normal((k, k)))
        return q * np.sign(np.diag(r))

    s = 10.0 ** (-12.0 * rng.uniform(0, 1, n_small))
    G = (haar(n_small) * s) @ haar(n_small).T
    t0 = time.time()
    _, _, _, info, stat = dgesvj(G, want_v=False)
    return {"n": n_small, "seconds": time.time() - t0, "sweeps": int(stat[3]), "info": int(info)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--alg", type=int, default=9)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--max-sweeps", type=int, default=80,
                    help="level-2 sweep cap (the reference's HYPJAC_NSWEEP is 30; this "
                         "matrix needs more at N=8192)")
    ap.add_argument("--form", choices=["learned+split", "learned", "split", "plain"],
                    default="learned+split",
                    help="form timed as `value`; plain is timed beside it unless --no-plain")
    ap.add_argument("--tau", type=float, default=1e-3,
                    help="split Gram while a pair's last max |cos| exceeds tau")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the profiled run (HIP events around every launch) that gives the "
                         "per-kernel breakdown")
    ap.add_argument("--flow", choices=["on", "off"], default="off",
                    help="dataflow sweeps (one persistent launch per sweep) for both forms")
    ap.add_argument("--staged", choices=["on", "off"], default="on",
                    help="staged sweeps (pair visits overlapped with half-step block updates)")
    ap.add_argument("--model", default=None, help="skip-model file (default: the packaged one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dgesvj", action="store_true", help="also time LAPACK dgesvj at N=1024")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU level "
              "(the reference's is not in its repository); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    os.environ.setdefault("HJ_QUIET", "1")
    import torch
    import gpujachx_amd as g

    n, alg = args.n, args.alg
    g.configureGPU(0)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    g.set_max_sweeps(args.max_sweeps)
    g.set_dataflow(args.flow == "on")
    g.set_staged(args.staged == "on")
    L = g.lib()
    G0, s_exact = make_input(torch, n, args.seed)
    W = torch.empty_like(G0)
    D = torch.empty(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    tm = (ctypes.c_double * 4)()
    model_info = None

    def set_form(form):
        nonlocal model_info
        bits = 0
        if "learned" in form:
            import gpujachx_amd.learned as lr
            _, thr, ms = lr.enable(args.model or lr.DEFAULT_MODEL)
            model_info = {"file": os.path.relpath(args.model or lr.DEFAULT_MODEL, ROOT),
                          "threshold": thr, "max_skips": ms}
            bits |= g.FORM_LEARNED_SKIP
        if "split" in form:
            bits |= g.FORM_SPLIT_GRAM
            g.set_split_tau(args.tau)
        g.set_form(bits)

    def one():
        W.copy_(G0)
        rc = L.hypJacL2_dev(alg, n, n, n, W.data_ptr(), n, None, 0, D.data_ptr(), ctypes.byref(sw),
                            ctypes.byref(gs), ctypes.byref(gb), tm, stream.cuda_stream)
        if rc:
            raise RuntimeError("hypJacL2_dev returned %d" % rc)

    def timed(form, steps, warmup, profile):
        """Wall time per SVD of `form` over `steps` runs after `warmup`; with
        profile, per-kernel HIP-event durations over the same runs."""
        set_form(form)
        for _ in range(warmup):
            one()
        g.set_profile(profile)
        kms, kcnt = np.zeros(6), np.zeros(6)
        visits = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            one()
            for k in range(6):
                ms, c = g.kernel_ms(k)
                kms[k] += ms
                kcnt[k] += c
            visits += g.last_skips()[2]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        g.set_profile(False)
        rot_s, rot_b, _ = g.last_sweeps()
        skips = g.last_skips()
        splits = g.last_splits()
        sig = torch.sqrt(torch.abs(D)).cpu().numpy()
        return {"per": elapsed / steps, "kms": kms, "kcnt": kcnt, "visits": visits,
                "flow": g.last_dataflow(), "staged": g.last_staged(),
                "sweeps": sw.value, "rot_b": rot_b, "skipped": skips[1], "split": splits[1],
                "split_per_sweep": splits[0], "skip_per_sweep": skips[0], "sigma": sig}

    other = "plain" if args.form != "plain" else "learned+split"
    # `value`: the SVDs as the library runs them, one captured graph per
    # sweep, with HIP events around each sweep's graph launch on the
    # library's stream (kind 5); a second run of the same SVD with HIP events
    # around every launch (step-by-step launches) gives the per-kernel
    # breakdown; the plain form is timed like `value`
    main_run = timed(args.form, args.steps, args.warmup, False)
    prof_run = None if args.no_profile else timed(args.form, 1, 0, True)
    cmp_run = None if args.no_plain else timed(other, 1, 0, False)
    g.set_form(g.FORM_PLAIN)
    per = main_run["per"]
    pk = prof_run if prof_run is not None else main_run
    kms, kcnt = pk["kms"], pk["kcnt"]
    sweeps = main_run["sweeps"]
    rot_b = main_run["rot_b"]
    sig = main_run["sigma"]

    # accuracy against the constructed spectrum
    ex = s_exact.cpu().numpy()
    sig_s, ex_s = np.sort(sig)[::-1], np.sort(ex)[::-1]
    rel = np.abs(sig_s - ex_s) / ex_s
    normwise = np.max(np.abs(sig_s - ex_s)) / ex_s[0]

    # roofline of the dominant kernel, the block update of a step fused with
    # the next step's Gram (k_upgram; the last step of a sweep runs the plain
    # update): algorithmic bytes per launch = executed pair visits per launch
    # x bytes per visit (read + write of the m x 32 pair: 512 m); flops per
    # visit 2048 m (update) + 1536 m (fused Gram).  The standalone Gram runs
    # once per sweep (step 0): 256 m bytes per visit.
    np_ = n // 32
    steps_per_sweep = n // 16 - 1
    pairs_per_launch = pk["visits"] / max(kcnt[2], 1)
    upd_bytes = pairs_per_launch * 512.0 * n
    gram_bytes = pairs_per_launch * 256.0 * n
    upd_ms = kms[2] / max(kcnt[2], 1)
    gram_ms = kms[0] / max(kcnt[0], 1)
    pair_ms = kms[1] / max(kcnt[1], 1)
    ach = upd_bytes / (upd_ms * 1e-3) / 1e9
    gram_ach = gram_bytes / (gram_ms * 1e-3) / 1e9
    upd_tf = pairs_per_launch * (2048.0 + 1536.0) * n / (upd_ms * 1e-3) / 1e12
    gram_tf = pairs_per_launch * 1536.0 * n / (gram_ms * 1e-3) / 1e12

    # HBM traffic per launch: FETCH_SIZE + WRITE_SIZE bytes per executed pair
    # visit of k_upgram, from counter runs of lib/hj_pmc on full-step launches
    # at N=8192 (each counter scaled by its reading for a copy of known size;
    # gpujachx_amd/pmc_traffic.py), times this run's executed visits per
    # launch, scaled linearly in the row count
    flow_used = bool(main_run["flow"])
    if flow_used:
        # the dataflow launch holds the fused transitions of steps 0..S-2 of
        # a sweep (the last step's update is a k_update after it) and every
        # pair visit; its roofline is the HBM stream of those transitions
        fl_n = max(kcnt[3], 1)
        pairs_per_launch = (pk["visits"] / fl_n) * (steps_per_sweep - 1) / steps_per_sweep
        upd_bytes = pairs_per_launch * 512.0 * n
        upd_ms = kms[3] / fl_n
        ach = upd_bytes / (upd_ms * 1e-3) / 1e9
        upd_tf = pairs_per_launch * (2048.0 + 1536.0) * n / (upd_ms * 1e-3) / 1e12
    staged_used = bool(main_run["staged"])
    timing_src = "profiled run: HIP events around every launch"
    if staged_used:
        # staged sweep: after step 0's Gram and visits, launches of half a
        # step's fused transitions each (beside pair visits), and the last
        # step's visits + plain update; together they update every executed
        # pair once (512 m bytes), and the visits' latency hides under them.
        # Timed in the `value` run itself: the sweeps' GPU time (events
        # around each graph launch) over their staged launches
        st_n = max(main_run["kcnt"][5], 1)
        pairs_per_launch = main_run["visits"] / st_n
        upd_bytes = pairs_per_launch * 512.0 * n
        upd_ms = main_run["kms"][5] / st_n
        timing_src = ("the timed region: HIP events around each sweep's graph launch, the sweeps' "
                      "GPU time over their k_stage launches (step 0's Gram and visits, the plan "
                      "and the last update, about 1 % of a sweep, included)")
        ach = upd_bytes / (upd_ms * 1e-3) / 1e9
        upd_tf = pairs_per_launch * (2048.0 + 1536.0) * n / (upd_ms * 1e-3) / 1e12
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "gpujachx_amd", "data", "upgram_traffic.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        up = tj["kernels"]["k_stage" if staged_used and "k_stage" in tj["kernels"] else "k_upgram"]
        traffic = up["bytes_per_visit"] * pairs_per_launch * n / tj["n"]
        traffic_src = {"file": "gpujachx_amd/data/upgram_traffic.json",
                       "kernel": "k_stage" if staged_used and "k_stage" in tj["kernels"] else "k_upgram",
                       "bytes_per_visit_at_n": [up["bytes_per_visit"], tj["n"]],
                       "traffic_over_algorithmic": up["traffic_over_algorithmic"],
                       "counter_scale": up.get("counter_scale", [tj["calibration"]["fetch_scale"],
                                                                 tj["calibration"]["write_scale"]])}

    out = {
        "metric": METRIC,
        "value": per,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if "split" not in args.form else "f64 (split Gram: bf16x3 MFMA, fp32 acc)",
        "data": "synthetic",
        "config": {"workload": "SVD M=N=%d, G=U diag(10^-12u) V^T (Haar), ALG %d, cycwor/cycwor, "
                               "%s form" % (n, alg, args.form),
                   "M": n, "N": n, "alg": alg, "form": args.form, "seed": args.seed,
                   "max_sweeps": args.max_sweeps, "tau": args.tau if "split" in args.form else None,
                   "skip_model": model_info if "learned" in args.form else None},
        "converged": bool(rot_b and rot_b[-1] == 0),
        "sweeps": sweeps,
        "visits_executed": main_run["visits"] // max(args.steps, 1),
        "visits_skipped": main_run["skipped"],
        "visits_split_gram": main_run["split"],
        "skip_per_sweep": main_run["skip_per_sweep"],
        "split_per_sweep": main_run["split_per_sweep"],
        "rot_b_per_sweep": rot_b,
        "sigma_err": {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
                      "normwise": float(normwise),
                      "frac_rel_le_1e-12": float(np.mean(rel <= 1e-12))},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": ("k_flow (dataflow sweep: fused block updates + Grams, pair visits "
                                "overlapped)") if flow_used else
                               ("k_stage (staged sweep: half a step's fused block updates + next-step "
                                "Grams beside pair visits)") if staged_used else
                               "k_upgram (block update + next-step Gram)",
                     "avg_launch_ms": upd_ms, "timing": timing_src,
                     "algorithmic_bytes_per_launch": upd_bytes,
                     "mfma_tflops": upd_tf, "mfma_frac_fp64": upd_tf / FP64_PEAK_TFS},
        "kernels": {"gram": {"avg_ms": gram_ms, "GBs": gram_ach, "tflops": gram_tf},
                    "pair": {"avg_ms": pair_ms},
                    "update": {"avg_ms": upd_ms, "GBs": ach, "tflops": upd_tf},
                    "stage": {"avg_ms": kms[4] / max(kcnt[4], 1)},
                    "launches_per_svd": float(main_run["kcnt"][5] / max(args.steps, 1) if staged_used
                                              else kcnt[3] if flow_used else kcnt[2]),
                    "source": "profiled run (HIP events around every launch)" if prof_run is not None
                              else "the timed run"},
        "dataflow": flow_used,
        "staged": staged_used,
        "profiled_s": ({"note": "the same SVD launched step by step with HIP events around "
                                "every launch (the per-kernel breakdown's run); `value` is the "
                                "library's graph-launched path",
                        args.form: prof_run["per"]} if prof_run is not None else None),
    }
    if cmp_run is not None:
        cs = np.sort(cmp_run["sigma"])[::-1]
        out[other] = {"s_per_svd": cmp_run["per"], "sweeps": cmp_run["sweeps"],
                      "visits_executed": cmp_run["visits"], "visits_skipped": cmp_run["skipped"],
                      "sigma_normwise_diff_vs_value": float(np.max(np.abs(cs - sig_s)) / sig_s[0])}
        out["speedup_vs_plain"] = (cmp_run["per"] / per) if other == "plain" else (per / cmp_run["per"])
        cmp_rel = np.abs(cs - ex_s) / ex_s
        out[other]["sigma_err"] = {"max_rel": float(cmp_rel.max()),
                                   "normwise": float(np.max(np.abs(cs - ex_s)) / ex_s[0])}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(G0.T.cpu().numpy().copy(), n, alg, sweeps,
                                           steps_per_sweep)
    if args.dgesvj:
        out["dgesvj_sample"] = dgesvj_sample()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
