"""Profiling driver: a few level-1 steps of the N x N SVD workload of bench.py,
preceded by a device-to-device copy of known size (the calibration kernel for
the FETCH_SIZE / WRITE_SIZE counters).

    rocprofv3 --kernel-trace --stats -d OUT -- python -m gpujachx_amd.profile_steps
    rocprofv3 --pmc FETCH_SIZE -d OUT -- python -m gpujachx_amd.profile_steps --steps 8
"""
import argparse
import ctypes
import os
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--alg", type=int, default=9)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--form", choices=["plain", "split"], default="plain",
                    help="split: every visit of the first sweep uses the split Gram")
    args = ap.parse_args()
    os.environ.setdefault("HJ_QUIET", "1")
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import torch
    import gpujachx_amd as g
    from bench import make_input

    n = args.n
    g.configureGPU(0)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    G0, _ = make_input(torch, n, 3)
    a = torch.empty(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)  # calibration: reads and writes copy_mib MiB
    torch.cuda.synchronize()
    D = torch.empty(n, dtype=torch.float64, device="cuda")
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    g.set_form(g.FORM_SPLIT_GRAM if args.form == "split" else g.FORM_PLAIN)
    g.set_step_limit(args.steps)
    rc = g.lib().hypJacL2_dev(args.alg, n, n, n, G0.data_ptr(), n, None, 0, D.data_ptr(),
                              ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), None,
                              torch.cuda.current_stream().cuda_stream)
    g.set_step_limit(0)
    g.set_form(g.FORM_PLAIN)
    torch.cuda.synchronize()
    print("profile_steps: n=%d steps=%d form=%s rc=%d copy=%d MiB"
          % (n, args.steps, args.form, rc, args.copy_mib), flush=True)


if __name__ == "__main__":
    main()
