"""x1 end to end on the GPU: HDF5 input in the reference's layout, the
reference's command line, stdout lines and output group, checked by the
verifier and against the oracle's rotation counts."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import svd_input
from test_tools import H5CHECK, H5GEN, have_tools, parse, stage

pytestmark = pytest.mark.gpu
X1 = os.path.join(g.LIB_DIR, "x1")


@pytest.mark.skipif(not have_tools, reason="HDF5 tools not built")
@pytest.mark.parametrize("alg", [9, 11])
def test_x1_end_to_end(tmp_path, alg):
    m = n = 256
    s = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    G = svd_input(m, n, s, seed=1)
    fin = stage(str(tmp_path), "in", G, n, lam=s ** 2)
    fres = os.path.join(str(tmp_path), "out.h5")
    env = dict(os.environ, HJ_QUIET="0")
    p = subprocess.run([X1, "0", g.STRAT_SO, "cycwor", "cycwor", str(alg), fin, str(n), fres],
                       capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    sweeps = [l for l in lines if l.startswith("BLK_SWP(")]
    assert sweeps and re.match(r"BLK_SWP\(\s*\d+\), ROT_S\(\s*\d+\), ROT_B\(\s*\d+\), TIME\(", sweeps[0])
    glb = re.search(r"GLB_ROT_S\(\s*(\d+)\), GLB_ROT_B\(\s*(\d+)\)", p.stdout)
    tot = re.search(r"s\s+(\d+) sweeps", p.stdout)
    assert glb and tot
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    s0, s1 = g.current_strats()
    ro = o.hypJacL2(alg, G, n, s0, s1)
    assert (int(glb.group(1)), int(glb.group(2))) == (ro["rot_s"], ro["rot_b"])
    assert int(tot.group(1)) == ro["sweeps"] == len(sweeps)
    q = subprocess.run([H5CHECK, fin, str(n), fres], capture_output=True, text=True, timeout=300)
    assert q.returncode == 0, q.stderr
    v = parse(q.stdout)
    assert v["|| U^T U - I ||_M"] < 1e-13
    assert abs(v["max rel err"]) < 1e-9
    if alg & 2:
        assert v["|| V^T J V - J ||_M"] < 1e-13


@pytest.mark.skipif(not have_tools, reason="HDF5 tools not built")
def test_x1_argument_error_exit_code(tmp_path):
    # nrow not a multiple of 64: hypJacL2 returns -2, x1 reports and returns it
    G = np.ones((96, 32))
    fin = stage(str(tmp_path), "in", G, 32)
    p = subprocess.run([X1, "0", g.STRAT_SO, "cycwor", "cycwor", "9", fin, "32"],
                       capture_output=True, text=True, timeout=120)
    assert "error -2" in p.stderr
    assert p.returncode == (256 - 2) % 256


@pytest.mark.skipif(not have_tools, reason="HDF5 tools not built")
@pytest.mark.parametrize("fact,ofact", [("qr", o.FACT_QR), ("drmac", o.FACT_QR_DRMAC)])
def test_x1_factorization_from_environment(tmp_path, fact, ofact):
    # HJ_FACT selects the reference's USE_QR / USE_DRMAC variants at run time
    m = n = 128
    s = 10.0 ** (-8.0 * np.arange(n) / (n - 1))
    G = svd_input(m, n, s, seed=2)
    fin = stage(str(tmp_path), "in", G, n, lam=s ** 2)
    env = dict(os.environ, HJ_QUIET="0", HJ_FACT=fact)
    p = subprocess.run([X1, "0", g.STRAT_SO, "cycwor", "cycwor", "9", fin, str(n)],
                       capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stderr
    glb = re.search(r"GLB_ROT_S\(\s*(\d+)\), GLB_ROT_B\(\s*(\d+)\)", p.stdout)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    s0, s1 = g.current_strats()
    old = o.set_factorization(ofact)
    try:
        ro = o.hypJacL2(9, G, n, s0, s1)
    finally:
        o.set_factorization(old)
    assert (int(glb.group(1)), int(glb.group(2))) == (ro["rot_s"], ro["rot_b"])
    bad = subprocess.run([X1, "0", g.STRAT_SO, "cycwor", "cycwor", "9", fin, str(n)],
                         capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, HJ_FACT="lu"))
    assert bad.returncode != 0 and "HJ_FACT" in bad.stderr
