"""Inputs of a counter run of a whole SVD in a fast form: writes the
headline matrix class (G = U diag(10^-12u) V^T, Haar U and V, seed 7,
column-major float64) and the packaged skip model (float32 parameters)
to files, for the Python-free driver `lib/hj_pmc learned` (rocprofv3's
counter collection fails host-side when the run is driven from Python).

    python -m gpujachx_amd.pmc_learned 8192 /tmp/G.bin /tmp/model.bin
    cd /tmp && export TMPDIR=/tmp
    timeout -s KILL 300 rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o pmc \\
        -- gpujachx_amd/lib/hj_pmc learned 8192 /tmp/G.bin /tmp/model.bin THRESHOLD MAX_SKIPS
    python -m gpujachx_amd.pmc_traffic FETCH.csv WRITE.csv --run-json RUN.json

Prints the threshold and the consecutive-skip cap to pass to the driver.
"""
import argparse

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("n", type=int)
    ap.add_argument("g_out")
    ap.add_argument("model_out")
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)
    import torch
    import gpujachx_amd.learned as lr

    n = args.n
    gen = torch.Generator(device="cuda").manual_seed(args.seed)

    def haar():
        q, r = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
        return q * torch.sign(torch.diagonal(r))

    s = torch.pow(10.0, -12.0 * torch.rand(n, dtype=torch.float64, device="cuda", generator=gen))
    G = (haar() * s) @ haar().T
    # column-major: the row-major transpose
    G.T.contiguous().cpu().numpy().tofile(args.g_out)
    p, thr, ms, _ = lr.load()
    np.asarray(p, dtype=np.float32).tofile(args.model_out)
    print("%r %d" % (float(thr), int(ms)))


if __name__ == "__main__":
    main()
