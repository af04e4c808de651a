"""HDF5 surface: h5gen writes the reference's input layout (IDADIM, G), the
verifier h5check restates var/pp80.f90's metrics.  CPU-only: the result
group is staged from the oracle's output."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input

H5GEN = os.path.join(g.LIB_DIR, "h5gen")
H5CHECK = os.path.join(g.LIB_DIR, "h5check")
have_tools = os.path.exists(H5GEN) and os.path.exists(H5CHECK)


def write_bin(path, a):
    np.ascontiguousarray(a, dtype="<f8").tofile(path)


def stage(tmp, name, G, nplus, lam=None, res=None):
    m, n = G.shape
    gb = os.path.join(tmp, name + "_G.bin")
    write_bin(gb, G.T)  # Fortran order = transpose in C order
    args = [H5GEN, os.path.join(tmp, name + ".h5"), str(n), str(m), str(m), str(n), str(nplus), gb]
    if lam is not None:
        lb = os.path.join(tmp, name + "_L.bin")
        write_bin(lb, lam)
        args.append(lb)
    if res is not None:
        U, V, D = res
        ub, vb, db = (os.path.join(tmp, name + s) for s in ("_U.bin", "_V.bin", "_D.bin"))
        write_bin(ub, U.T)
        write_bin(vb, V.T)
        write_bin(db, D)
        args[7] = ub
        if lam is not None:
            args.pop()
        args += [vb, db]
    subprocess.check_call(args)
    return os.path.join(tmp, name + ".h5")


def parse(out):
    vals = {}
    for line in out.splitlines():
        m = re.match(r"(.+?)\s+([-+0-9.eE]+)$", line.strip())
        if m:
            vals[m.group(1)] = float(m.group(2))
    return vals


@pytest.mark.skipif(not have_tools, reason="HDF5 tools not built")
@pytest.mark.parametrize("nplus", [128, 40])
def test_h5check_on_oracle_result(tmp_path, nplus):
    m, n = 256, 128
    if nplus == n:
        s = 10.0 ** (-5.0 * np.arange(n) / (n - 1))
        G = svd_input(m, n, s, seed=5)
        lam = s ** 2
    else:
        G = graded_input(m, n, seed=6, lo=-2, hi=2)
        lam = None
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // 16)
    r = o.hypJacL2(11, G, nplus, s0, s1, nthreads=2)
    fin = stage(str(tmp_path), "in", G, nplus, lam=lam)
    fres = stage(str(tmp_path), "res", G, nplus, res=(r["G"], r["V"], r["D"]))
    p = subprocess.run([H5CHECK, fin, str(n), fres], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    v = parse(p.stdout)
    assert v["|| U^T U - I ||_M"] < 1e-13
    assert v["|| V^T J V - J ||_M"] < 1e-10
    # column-relative residual is eps*||G||/sigma_j at worst (kappa = 1e5 here)
    assert v["|| G V - U S ||_M"] < 1e-13 * np.abs(G).max() * n
    assert v["|| G V - U S ||_R"] < 1e-9
    if lam is not None:
        assert abs(v["max rel err"]) < 1e-9


@pytest.mark.skipif(not have_tools, reason="HDF5 tools not built")
def test_x1_usage_message():
    x1 = os.path.join(g.LIB_DIR, "x1")
    p = subprocess.run([x1], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert "DEV SDY SNP0 SNP1 ALG H5F H5G [H5R]" in p.stderr
