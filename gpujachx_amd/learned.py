"""Learned pair-skip: traces of the plain method, training set, GPU training
with the library's own HIP kernels, and the model file.

A visit of a block pair (one level-1 step of one pair) is *skippable* when
it would perform no big rotation.  The model sees the pair's record from its
last executed visit (rotation counts, the largest |cos| between two of its
columns before rotating, the two block norms) plus run-level context (sweep
index, the previous sweep's big-rotation fraction, J-sign mix, problem size)
and outputs P(big rotation).  The library skips visits with P below a
threshold, never more than `max_skips` in a row, and always ends a run with
an unskipped sweep that passes the reference's test.

    python -m gpujachx_amd.learned train --out gpujachx_amd/data/skip_mlp.json
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

import gpujachx_amd as g

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "skip_mlp.json")
REC_FIELDS = ("rot_s", "rot_b", "maxcos", "l2ni", "l2nj", "sweep", "skips", "pad")
REC_DTYPE = np.dtype([("rot_s", "<u4"), ("rot_b", "<u4"), ("maxcos", "<f4"), ("l2ni", "<f4"),
                      ("l2nj", "<f4"), ("sweep", "<u4"), ("skips", "<u4"), ("pad", "<u4")])


def nparam():
    return g.lib().hj_mlp_nparam()


def nfeat():
    return g.lib().hj_mlp_nfeat()


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def init_params(seed=0, hidden=64):
    """He-initialised parameter vector in the library's layout."""
    F, H = nfeat(), hidden
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(F * H) * np.sqrt(2.0 / F), np.zeros(H),
             rng.standard_normal(H * H) * np.sqrt(2.0 / H), np.zeros(H),
             rng.standard_normal(H) * np.sqrt(1.0 / H), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == nparam()
    return p


def collect_trace(alg, G, nplus, max_sweeps=80):
    """Run the plain method on G (host array) and return (records
    [sweeps, keys] of REC_DTYPE, strat1 [steps, np, 2], per-sweep rot_b)."""
    n = G.shape[1]
    g.set_form(g.FORM_PLAIN)
    s0, s1 = g.current_strats()
    keys = s1.shape[0] * s1.shape[1]
    assert g.lib().hj_visit_record_size() == REC_DTYPE.itemsize
    buf = np.zeros((max_sweeps, keys), dtype=REC_DTYPE)
    g.lib().hj_set_visit_trace(buf.ctypes.data, max_sweeps)
    old = g.set_max_sweeps(max_sweeps)
    try:
        r = g.hypJacL2(alg, G, nplus)
    finally:
        g.lib().hj_set_visit_trace(None, 0)
        g.set_max_sweeps(old)
    _, rot_b, _ = g.last_sweeps()
    return buf[:r.sweeps].copy(), s1, np.array(rot_b, dtype=np.float64), r


def features(rec, strat1, nplus, n1, sweep_now, prev_frac):
    """Features of the next visit for every record (device code)."""
    rec = np.ascontiguousarray(rec, dtype=REC_DTYPE)
    s1 = np.ascontiguousarray(strat1, dtype=np.uint16)
    X = np.zeros((rec.size, nfeat()), dtype=np.float32)
    rc = g.lib().hj_skip_features(rec.ctypes.data, rec.size,
                                  s1.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)), nplus, n1,
                                  sweep_now, float(prev_frac), _pf(X))
    if rc:
        raise RuntimeError("hj_skip_features failed")
    return X


def dataset(trace, strat1, rot_b, nplus, n1):
    """(X, y): for every sweep k and pair, the features after sweep k and
    whether the pair's visit in sweep k+1 had a big rotation."""
    keys = trace.shape[1]
    max_rot = keys * 496.0
    Xs, ys = [], []
    for k in range(trace.shape[0] - 1):
        Xs.append(features(trace[k], strat1, nplus, n1, k + 2, rot_b[k] / max_rot))
        ys.append((trace[k + 1]["rot_b"] > 0).astype(np.float32))
    return np.concatenate(Xs), np.concatenate(ys)


def train(X, y, params=None, epochs=6, batch=4096, lr=2e-3, wpos=4.0, seed=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    p = init_params(seed) if params is None else np.array(params, dtype=np.float32)
    losses = np.zeros(epochs, dtype=np.float32)
    rc = g.lib().hj_mlp_train(_pf(X), _pf(y), X.shape[0], epochs, batch, lr, wpos, seed, _pf(p),
                              _pf(losses))
    if rc:
        raise RuntimeError("hj_mlp_train failed")
    return p, losses


def predict(X, params):
    X = np.ascontiguousarray(X, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(X.shape[0], dtype=np.float32)
    if g.lib().hj_mlp_predict(_pf(X), X.shape[0], _pf(p), _pf(out)):
        raise RuntimeError("hj_mlp_predict failed")
    return out


def grad(X, y, params, wpos=4.0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    gr = np.zeros_like(p)
    loss = np.zeros(1, dtype=np.float32)
    if g.lib().hj_mlp_grad(_pf(X), _pf(y), X.shape[0], _pf(p), wpos, _pf(gr), _pf(loss)):
        raise RuntimeError("hj_mlp_grad failed")
    return gr, float(loss[0])


def save(path, params, threshold, max_skips, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    json.dump({"nfeat": nfeat(), "hidden": 64, "threshold": threshold, "max_skips": max_skips,
               "params": [float(v) for v in np.asarray(params, dtype=np.float32)], "meta": meta},
              open(path, "w"))


def load(path=DEFAULT_MODEL):
    d = json.load(open(path))
    return np.array(d["params"], dtype=np.float32), d["threshold"], d["max_skips"], d


def enable(path=DEFAULT_MODEL, threshold=None, max_skips=None, form=None):
    """Load the model into the library and switch the learned form on."""
    p, thr, ms, _ = load(path)
    thr = thr if threshold is None else threshold
    ms = ms if max_skips is None else max_skips
    if g.lib().hj_set_skip_model(_pf(p), p.size, thr, ms):
        raise RuntimeError("model size mismatch")
    g.set_form(g.FORM_LEARNED_SKIP if form is None else form)
    return p, thr, ms


def _configs():
    """Training inputs: the plain method's traces on SVD 256, HSVD
    4096x2048, SVD 2048, SVD 8192 and HSVD 8192 (NPLUS = N/2), with seeds
    other than the benchmarks'."""
    return [
        ("svd 256", 256, 256, 256, 101),
        ("hsvd 4096x2048", 4096, 2048, 1024, 102),
        ("svd 2048", 2048, 2048, 2048, 105),
        ("svd 8192", 8192, 8192, 8192, 103),
        ("hsvd 8192", 8192, 8192, 4096, 104),
    ]


def balance(Xs, ys, per_config, rng):
    """Equal weight per configuration: each trace contributes per_config
    samples (subsampled, or repeated when it has fewer)."""
    Xb, yb = [], []
    for X, y in zip(Xs, ys):
        idx = rng.choice(X.shape[0], per_config, replace=X.shape[0] < per_config)
        Xb.append(X[idx])
        yb.append(y[idx])
    return np.concatenate(Xb), np.concatenate(yb)


def _make_input(torch, m, n, nplus, seed, kind):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    dev = torch.device("cuda")
    if kind == "hsvd":
        Gt = torch.randn(n, m, dtype=torch.float64, device=dev, generator=gen)
        sc = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device=dev, generator=gen) * 8 - 4)
        return (Gt * sc[:, None]).cpu().numpy().T.copy()

    def haar(k):
        a = torch.randn(k, k, dtype=torch.float64, device=dev, generator=gen)
        q, r = torch.linalg.qr(a)
        return q * torch.sign(torch.diagonal(r))
    U, V = haar(m)[:, :n], haar(n)
    if n == 256:
        s = torch.pow(10.0, -6.0 * torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    else:
        s = torch.pow(10.0, -12.0 * torch.rand(n, dtype=torch.float64, device=dev, generator=gen))
    return ((U * s) @ V.T).cpu().numpy()


def _run_dev(torch, alg, Gt, nplus, D):
    """One hypJacL2_dev call on a fresh copy of the column-major device
    matrix Gt (n x m row-major view); returns wall seconds."""
    m, n = Gt.shape[1], Gt.shape[0]
    W = Gt.clone()
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    tm = (ctypes.c_double * 4)()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = g.lib().hypJacL2_dev(alg, m, n, nplus, W.data_ptr(), m, None, 0, D.data_ptr(), ctypes.byref(sw),
                              ctypes.byref(gs), ctypes.byref(gb), tm,
                              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    if rc:
        raise RuntimeError("hypJacL2_dev returned %d" % rc)
    return time.perf_counter() - t0


def evaluate(path, seeds=(201,), sizes=None, thresholds=(None,), max_skips=(None,), max_sweeps=80):
    """Plain vs learned form (device API, warm) on fresh matrices of the
    training shapes, for each (threshold, max_skips) setting."""
    import torch
    out = []
    for name, m, n, nplus, _ in _configs():
        if sizes and n not in sizes:
            continue
        for seed in seeds:
            Gt = torch.from_numpy(_make_input(torch, m, n, nplus, seed,
                                              "hsvd" if nplus < n else "svd").T.copy()).cuda()
            D = torch.empty(n, dtype=torch.float64, device="cuda")
            alg = 11 if (nplus < n and m != n) else 9
            g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
            old = g.set_max_sweeps(max_sweeps)
            try:
                g.set_form(g.FORM_PLAIN)
                _run_dev(torch, alg, Gt, nplus, D)
                tp = _run_dev(torch, alg, Gt, nplus, D)
                sp = np.sort(torch.sqrt(torch.abs(D)).cpu().numpy())
                swp, vp = len(g.last_sweeps()[0]), g.last_skips()[2]
                for thr in thresholds:
                    for ms in max_skips:
                        enable(path, threshold=thr, max_skips=ms)
                        tl = _run_dev(torch, alg, Gt, nplus, D)
                        sk, tot, vl = g.last_skips()
                        _, rb, _ = g.last_sweeps()
                        sl = np.sort(torch.sqrt(torch.abs(D)).cpu().numpy())
                        rec = {"config": name, "seed": seed, "threshold": thr, "max_skips": ms,
                               "plain_s": tp, "learned_s": tl, "speedup": tp / tl,
                               "plain_sweeps": swp, "learned_sweeps": len(rb),
                               "plain_visits": vp, "learned_visits": vl, "skipped": tot,
                               "skipped_per_sweep": sk, "rot_b_per_sweep": rb,
                               "ends_unskipped_converged": bool(sk[-1] == 0 and rb[-1] == 0),
                               "sigma_normwise_diff": float(np.abs(sp - sl).max() / sp.max())}
                        print(json.dumps(rec), flush=True)
                        out.append(rec)
                        g.set_form(g.FORM_PLAIN)
            finally:
                g.set_form(g.FORM_PLAIN)
                g.set_max_sweeps(old)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--max-samples", type=int, default=3_000_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--max-skips", type=int, default=8)
    ap.add_argument("--max-n", type=int, default=8192, help="largest training order")
    ap.add_argument("--sizes", type=int, nargs="*", help="eval: orders to evaluate")
    ap.add_argument("--thresholds", type=float, nargs="*", help="eval: skip thresholds")
    ap.add_argument("--skips", type=int, nargs="*", help="eval: max consecutive skips")
    args = ap.parse_args()
    os.environ.setdefault("HJ_QUIET", "1")
    import torch
    g.configureGPU(0)
    if args.cmd == "eval":
        evaluate(args.model, sizes=args.sizes, thresholds=args.thresholds or (None,),
                 max_skips=args.skips or (None,))
        return
    Xs, ys, meta = [], [], {"configs": []}
    for name, m, n, nplus, seed in _configs():
        if n > args.max_n:
            continue
        t0 = time.time()
        G = _make_input(torch, m, n, nplus, seed, "hsvd" if (nplus < n and m != n) else "svd")
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        tr, s1, rb, r = collect_trace(11 if (nplus < n and m != n) else 9, G, nplus)
        X, y = dataset(tr, s1, rb, nplus, n // 16)
        meta["configs"].append({"name": name, "sweeps": int(r.sweeps), "samples": int(X.shape[0]),
                                "positives": float(y.mean()), "seconds": time.time() - t0})
        print(name, meta["configs"][-1], flush=True)
        Xs.append(X)
        ys.append(y)
    rng = np.random.default_rng(0)
    X, y = balance(Xs, ys, args.max_samples // len(Xs), rng)
    perm = rng.permutation(X.shape[0])
    nval = max(1, X.shape[0] // 10)
    val, trn = perm[:nval], perm[nval:]
    t0 = time.time()
    params, losses = train(X[trn], y[trn], epochs=args.epochs)
    pv = predict(X[val], params)
    yv = y[val]
    skip = pv < args.threshold
    meta.update({"train_seconds": time.time() - t0, "epoch_loss": [float(v) for v in losses],
                 "val_samples": int(nval), "val_skip_rate": float(skip.mean()),
                 "val_false_skip_rate": float((skip & (yv > 0.5)).sum() / max(1, (yv > 0.5).sum())),
                 "val_recall_of_skippable": float((skip & (yv < 0.5)).sum() / max(1, (yv < 0.5).sum()))})
    print(json.dumps({k: v for k, v in meta.items() if k != "configs"}), flush=True)
    save(args.out, params, args.threshold, args.max_skips, meta)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
