"""HBM bytes per launch of the fused update kernel from two counter runs of
the profiling driver `lib/hj_pmc` (one FETCH_SIZE pass, one WRITE_SIZE pass,
rocprofv3 CSV output).  Each counter is scaled by what it reads for the
driver's calibration copy (a device-to-device copy of known size: its reads
and its writes are each COPY bytes), so a counter that under-counts by a
fixed factor on this GPU is corrected by that factor.

    cd /tmp && export TMPDIR=/tmp
    timeout -s KILL 90 rocprofv3 --pmc FETCH_SIZE --output-format csv -d /tmp/pmc_FETCH_SIZE -o pmc \
        -- <repo>/gpujachx_amd/lib/hj_pmc 8192 8
    (the same with WRITE_SIZE)
    python -m gpujachx_amd.pmc_traffic FETCH.csv WRITE.csv [--copy-mib 1024] [--out FILE]
    python -m gpujachx_amd.pmc_traffic FETCH.csv WRITE.csv --run-json RUN.json   (pmc_learned)

Writes a JSON file with bytes per launch and per executed pair visit
(N x 32 columns read and written) for k_upgram, k_update and k_gram; with
a counter run of one staged sweep (`hj_pmc 8192 0`), also for k_stage,
whose launches each update half a step's pairs (the few launches with
visits only are left out of its average).
"""
import argparse
import csv
import json
import os

KERNELS = {"k_upgram": "hj::k_upgram(", "k_update": "hj::k_update(", "k_gram": "hj::k_gram(",
           "k_pair": "hj::k_pair", "k_stage": "hj::k_stage<"}


def per_kernel(path):
    rows = list(csv.DictReader(open(path)))
    out = {"copy": None}
    for r in rows:
        name, v = r["Kernel_Name"], float(r["Counter_Value"])
        if name.startswith("__amd_rocclr_copyBuffer"):
            # the calibration copy is the largest copy dispatch
            out["copy"] = v if out["copy"] is None else max(out["copy"], v)
            continue
        for k, pre in KERNELS.items():
            if pre in name:
                out.setdefault(k, []).append(v)
    return out


def run_bytes(path, scale):
    """Counter total (KiB x scale) over the sweeps' kernels of a whole run
    (every hj:: dispatch but the final column normalisation), and the
    k_stage dispatches' values."""
    tot, stage = 0.0, []
    for r in csv.DictReader(open(path)):
        name = r["Kernel_Name"]
        if "hj::" not in name or "k_initD" in name:
            continue
        v = float(r["Counter_Value"]) * scale * 1024.0
        tot += v
        if "k_stage<" in name:
            stage.append(v)
    return tot, stage


def whole_run(args, fs, ws):
    run = json.loads(open(args.run_json).read().strip().splitlines()[-1])
    fb, fstage = run_bytes(args.fetch_csv, fs)
    wb, wstage = run_bytes(args.write_csv, ws)
    v = run["visits_executed"]
    ent = {"form": run["form"], "sweeps": run["sweeps"], "visits_executed": v,
           "visits_skipped": run["visits_skipped"], "fetch_bytes": fb, "write_bytes": wb,
           "bytes_per_visit": (fb + wb) / v,
           "traffic_over_algorithmic": (fb + wb) / v / (768.0 * run["n"]),
           "stage_dispatches": len(fstage),
           "stage_bytes_per_dispatch": (sum(fstage) + sum(wstage)) / max(len(fstage), 1),
           "counter_scale": [fs, ws],
           "note": "every hj:: dispatch of the run but the final normalisation, over the run's "
                   "executed visits; algorithmic 768 m B per visit"}
    out = json.load(open(args.out)) if os.path.exists(args.out) else {"n": run["n"], "kernels": {}}
    out.setdefault("runs", {})[run["form"]] = ent
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(ent))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--merge", action="store_true",
                    help="keep the other kernels' entries of an existing output file")
    ap.add_argument("--run-json", default=None,
                    help="the JSON line of a whole-run driver (pmc_learned): record the run's "
                         "bytes per executed visit under its form instead of per-launch figures")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                  "upgram_traffic.json"))
    args = ap.parse_args(argv)
    f, w = per_kernel(args.fetch_csv), per_kernel(args.write_csv)
    copy_kib = args.copy_mib * 1024.0
    fs, ws = copy_kib / f["copy"], copy_kib / w["copy"]
    if args.run_json:
        return whole_run(args, fs, ws)
    visits = args.n // 32            # pairs per step when none is skipped
    res = {"n": args.n, "visits_per_launch": visits,
           "calibration": {"copy_mib": args.copy_mib, "fetch_kib": f["copy"], "write_kib": w["copy"],
                           "fetch_scale": fs, "write_scale": ws},
           "kernels": {}}
    for k in KERNELS:
        if k not in f or k not in w:
            continue
        fl, wl = f[k], w[k]
        nv = visits
        if k == "k_stage":
            # launches with fused items (half a step's pairs each), not the
            # visits-only launches
            mf, mw = sorted(fl)[len(fl) // 2], sorted(wl)[len(wl) // 2]
            fl = [v for v in fl if v > 0.5 * mf]
            wl = [v for v in wl if v > 0.5 * mw]
            nv = visits // 2
        fk = sum(fl) / len(fl)
        wk = sum(wl) / len(wl)
        b = (fk * fs + wk * ws) * 1024.0
        alg = {"k_upgram": 512.0, "k_update": 512.0, "k_gram": 256.0, "k_stage": 512.0}.get(k)
        ent = {"dispatches": len(fl), "fetch_kib": fk, "write_kib": wk,
               "bytes_per_launch": b, "bytes_per_visit": b / nv, "counter_scale": [fs, ws]}
        if alg is not None:
            ent["algorithmic_bytes_per_launch"] = alg * args.n * nv
            ent["traffic_over_algorithmic"] = b / (alg * args.n * nv)
        res["kernels"][k] = ent
    if args.merge and os.path.exists(args.out):
        old = json.load(open(args.out))
        if old.get("n") == args.n:
            res["kernels"] = {**old["kernels"], **res["kernels"]}
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
