"""Seconds per (H)SVD and accuracy on the benchmark configurations other
than the headline one, for the plain FP64 form and the learned+split form.

  c1      SVD, M=N=256, s_i = 10^(-6(i-1)/255), Haar U, V; ALG 11 (V kept so
          that ||V^T V - I|| can be reported), cycwor/cycwor
  c2      HSVD, M=4096, N=2048, NPLUS=1024, G iid N(0,1) x column scales
          10^U(-4,4); ALG 11
  c3      SVD, M=N=8192, G = U diag(s) V^T, s log-uniform in [1e-12, 1];
          ALG 9 (the headline matrix class, another seed)
  c5svd   SVD, M=N=16384, G = U diag(s) V^T diag(10^u), s log-uniform in
          [1e-10, 1], column grading u_j ~ U(-4, 4) as in c2; ALG 9
  c5hsvd  the same matrix with NPLUS = N/2; ALG 9

    python -m gpujachx_amd.config_bench c1 c2 [--forms plain learned+split plain32]

(plain32 / learned32: the plain FP64 method / the learned pair-skip on
32-column blocks, 64-column pairs.)

One JSON line per (configuration, form): s per call (device-resident
input, after one warm-up call for the small configurations), sweeps,
executed/skipped visits, and the accuracy measures the configuration names
(sigma against the constructed spectrum where G has it; ||V^T J V - J||; the
residual ||G V - U diag(sqrt|D|)|| / ||G||); every form after the first is
also compared per value with the first (plain FP64) on the same matrix.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from gpujachx_amd.accuracy import per_sigma

CONFIGS = {
    "c1": dict(m=256, n=256, nplus=256, alg=11, kind="svd_geom", warm=True),
    "c3": dict(m=8192, n=8192, nplus=8192, alg=9, kind="svd_log12", warm=False),
    "c2": dict(m=4096, n=2048, nplus=1024, alg=11, kind="graded", warm=True),
    "c5svd": dict(m=16384, n=16384, nplus=16384, alg=9, kind="svd_log10_graded", warm=False),
    "c5hsvd": dict(m=16384, n=16384, nplus=8192, alg=9, kind="svd_log10_graded", warm=False),
}


def make(torch, cfg, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    dev = torch.device("cuda")
    m, n = cfg["m"], cfg["n"]

    def haar(k):
        a = torch.randn(k, k, dtype=torch.float64, device=dev, generator=gen)
        q, r = torch.linalg.qr(a)
        return q * torch.sign(torch.diagonal(r))

    s = None
    if cfg["kind"] == "graded":
        Gt = torch.randn(n, m, dtype=torch.float64, device=dev, generator=gen)
        Gt.mul_(torch.pow(10.0, torch.rand(n, 1, dtype=torch.float64, device=dev, generator=gen) * 8 - 4))
        return Gt, None
    if cfg["kind"] == "svd_geom":
        s = torch.pow(10.0, -6.0 * torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    elif cfg["kind"] == "svd_log12":
        s = torch.pow(10.0, -12.0 * torch.rand(n, dtype=torch.float64, device=dev, generator=gen))
    else:
        s = torch.pow(10.0, -10.0 * torch.rand(n, dtype=torch.float64, device=dev, generator=gen))
    U = haar(m)[:, :n]
    V = haar(n)
    G = (U * s) @ V.T
    if cfg["kind"].endswith("_graded"):
        # column scales 10^U(-4,4): the spectrum of G is no longer s
        G = G * torch.pow(10.0, torch.rand(n, dtype=torch.float64, device=dev, generator=gen) * 8 - 4)
        return G.T.contiguous(), None
    return G.T.contiguous(), s


def run(g, torch, cfg, Gt, form, model):
    n, m, alg = cfg["n"], cfg["m"], cfg["alg"]
    bits = 0
    if "learned" in form:
        import gpujachx_amd.learned as lr
        lr.enable(model or lr.DEFAULT_MODEL)
        bits |= g.FORM_LEARNED_SKIP
    if "split" in form:
        bits |= g.FORM_SPLIT_GRAM
    g.set_form(bits)
    W = torch.empty_like(Gt)
    V = torch.empty(n, n, dtype=torch.float64, device="cuda") if alg & 2 else None
    D = torch.empty(n, dtype=torch.float64, device="cuda")
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    tm = (ctypes.c_double * 4)()

    def one():
        W.copy_(Gt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = g.lib().hypJacL2_dev(alg, m, n, cfg["nplus"], W.data_ptr(), m,
                                  V.data_ptr() if V is not None else None, n if V is not None else 0,
                                  D.data_ptr(), ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), tm,
                                  torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        if rc:
            raise RuntimeError("hypJacL2_dev returned %d" % rc)
        return time.perf_counter() - t0

    if cfg["warm"]:
        one()
    t = one()
    _, rb, _ = g.last_sweeps()
    sk = g.last_skips()
    g.set_form(g.FORM_PLAIN)
    return t, W, V, D, sw.value, rb, sk


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="+", choices=sorted(CONFIGS))
    ap.add_argument("--forms", nargs="+", default=["plain", "learned+split"])
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--max-sweeps", type=int, default=80)
    ap.add_argument("--model", default=None)
    ap.add_argument("--skip-min-frac", type=float, default=None)
    ap.add_argument("--skip-floor-mixed", type=float, default=None,
                    help="floor multiple for runs that mix J signs (default 3, 6 at width 32)")
    args = ap.parse_args()
    os.environ.setdefault("HJ_QUIET", "1")
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import torch
    import gpujachx_amd as g
    g.configureGPU(0)
    g.set_max_sweeps(args.max_sweeps)
    if args.skip_min_frac is not None:
        g.lib().hj_set_skip_min_frac(args.skip_min_frac)
    if args.skip_floor_mixed is not None:
        g.lib().hj_set_skip_floor_mixed(args.skip_floor_mixed)
    for name in args.configs:
        cfg = CONFIGS[name]
        Gt, s = make(torch, cfg, args.seed)
        g.init_strats_builtin("cycwor", 32, "cycwor", cfg["n"] // 16)
        base = None
        for form in args.forms:
            # plain32 / learned32: the plain FP64 method / the learned pair-skip
            # on 32-column blocks
            width = 32 if form.endswith("32") else 16
            g.lib().hj_set_block_width(width)
            if args.skip_floor_mixed is not None:
                g.lib().hj_set_skip_floor_mixed(args.skip_floor_mixed)
            try:
                t, W, V, D, sweeps, rb, sk = run(g, torch, cfg, Gt, form[:-2] if width == 32 else form,
                                                 args.model)
            finally:
                g.lib().hj_set_block_width(16)
            sig = torch.sqrt(D.abs())
            out = {"config": name, "form": form, "M": cfg["m"], "N": cfg["n"], "NPLUS": cfg["nplus"],
                   "alg": cfg["alg"], "s_per_call": t, "sweeps": sweeps,
                   "converged": bool(rb and rb[-1] == 0), "visits_executed": sk[2],
                   "visits_skipped": sk[1]}
            if s is not None and cfg["nplus"] == cfg["n"]:
                # the constructed spectrum is the singular values only for J = I
                a = torch.sort(sig, descending=True).values
                e = torch.sort(s, descending=True).values
                rel = ((a - e).abs() / e)
                out["sigma_err"] = {"normwise": float(((a - e).abs().max() / e[0]).item()),
                                    "max_rel": float(rel.max().item()),
                                    "frac_rel_le_1e-12": float((rel <= 1e-12).double().mean().item())}
            if V is not None:
                n, nplus = cfg["n"], cfg["nplus"]
                J = torch.cat([torch.ones(nplus, dtype=torch.float64, device="cuda"),
                               -torch.ones(n - nplus, dtype=torch.float64, device="cuda")])
                Vm, Gm, Um = V.T, Gt.T, W.T
                out["VtJV_minus_J_max"] = float((Vm.T @ (J[:, None] * Vm) - torch.diag(J)).abs().max().item())
                out["residual_rel"] = float(((Gm @ Vm - Um * sig).abs().max() / Gm.abs().max()).item())
            Dh = D.cpu().numpy().copy()
            if base is None:
                base = (form, Dh, t)
            else:
                out["per_sigma_vs_" + base[0]] = per_sigma(Dh, base[1])
                out["speedup_vs_" + base[0]] = base[2] / t
            print(json.dumps(out), flush=True)
            del W, V, D
            torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
