"""Strategy tables: the library's generator and loader against the
reference's own tables (oracle/_ref/strat.so, built from the table headers
under /root/reference/strat) and against the fingerprints recorded from
them (tests/golden/strat_fingerprints.json)."""
import ctypes
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
REF = o.REF_STRAT_SO
have_ref = os.path.exists(REF)


def ref_table(fam, n):
    L = ctypes.CDLL(REF)
    steps = n if fam == "mmstep" else n - 1
    name = "%s%05u" % (fam, n)
    try:
        addr = ctypes.c_void_p.in_dll(L, name)
    except ValueError:
        return None
    ct = ctypes.c_ubyte if n <= 256 else ctypes.c_ushort
    arr = (ct * (steps * (n // 2) * 2)).in_dll(L, name)
    return np.ctypeslib.as_array(arr).reshape(steps, n // 2, 2).astype(np.uint16)


def fingerprint(t):
    return hashlib.sha256(np.ascontiguousarray(t, dtype="<u2").tobytes()).hexdigest()[:16]


def generable():
    out = []
    for fam in g.FAMILIES:
        for n in [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]:
            out.append((fam, n))
        if fam in ("BrentL", "mmstep"):
            for n in range(48, 1024 + 1, 16):
                if n & (n - 1):
                    out.append((fam, n))
    return out


def valid(t, n, cyclic):
    for st in t:
        flat = st.reshape(-1)
        assert sorted(flat.tolist()) == list(range(n)), "step not a perfect matching"
        assert np.all(st[:, 0] < st[:, 1])
    pairs = [tuple(p) for st in t for p in st]
    if cyclic:
        assert len(set(pairs)) == len(pairs) == n * (n - 1) // 2
    else:
        assert set(pairs) == {(i, j) for i in range(n) for j in range(i + 1, n)}


def test_fingerprints():
    fp = json.load(open(os.path.join(GOLD, "strat_fingerprints.json")))
    assert len(fp) == 24
    for name, (S, P, h) in fp.items():
        fam, n = name[:6], int(name[6:])
        t = g.strat_generate(fam, n)
        assert t.shape == (S, P, 2)
        assert fingerprint(t) == h, name


@pytest.mark.parametrize("fam", g.FAMILIES)
def test_generated_tables_valid(fam):
    for n in [2, 4, 16, 32, 64, 256, 1024] + ([12, 48, 100, 960] if fam in ("BrentL", "mmstep") else []):
        t = g.strat_generate(fam, n)
        valid(t, n, fam != "mmstep")


@pytest.mark.skipif(not have_ref, reason="reference strat.so not built (no /root/reference)")
def test_generator_equals_reference_tables():
    checked = 0
    for fam, n in generable():
        r = ref_table(fam, n)
        if r is None:
            continue
        np.testing.assert_array_equal(g.strat_generate(fam, n), r, err_msg="%s %d" % (fam, n))
        checked += 1
    assert checked >= 80


def test_no_construction_raises():
    with pytest.raises(ValueError):
        g.strat_generate("cycwor", 12)
    with pytest.raises(ValueError):
        g.strat_generate("nosuch", 16)


def test_init_strats_from_plugin_matches_builtin():
    g.init_strats(g.STRAT_SO, "cycwor", 32, "cycwor", 512)
    a0, a1 = g.current_strats()
    assert g.strat_ids() == (6, 6)
    g.init_strats_builtin("cycwor", 32, "cycwor", 512)
    b0, b1 = g.current_strats()
    np.testing.assert_array_equal(a0, b0)
    np.testing.assert_array_equal(a1, b1)
    g.init_strats(g.STRAT_SO, "mmstep", 32, "BrentL", 48)
    a0, a1 = g.current_strats()
    assert a0.shape == (32, 16, 2) and a1.shape == (47, 24, 2)
    assert g.strat_ids() == (1, 2)


@pytest.mark.skipif(not have_ref, reason="reference strat.so not built")
def test_init_strats_reads_reference_plugin():
    # orders this library cannot construct still load from the reference's plugin
    g.init_strats(REF, "cycwor", 32, "cycwor", 12)
    _, t1 = g.current_strats()
    np.testing.assert_array_equal(t1, ref_table("cycwor", 12))
    g.init_strats(REF, "rowcyc", 32, "colcyc", 896)
    _, t1 = g.current_strats()
    np.testing.assert_array_equal(t1, ref_table("colcyc", 896))


@pytest.mark.parametrize("args,msg", [
    (("", "cycwor", 32, "cycwor", 16), "SDY empty"),
    (("S", "", 32, "cycwor", 16), "SNP0 empty"),
    (("S", "cycwor", 31, "cycwor", 16), "n0 not positive even"),
    (("S", "cycwor", 32, "", 16), "SNP1 empty"),
    (("S", "cycwor", 32, "cycwor", 0), "n1 not positive even"),
    (("S", "foo", 32, "cycwor", 16), "SNP0 unknown"),
    (("S", "cycwor", 32, "bar", 16), "SNP1 unknown"),
    (("S", "cycwor", 32, "cycwor", 12), "NO SUCH STRATEGY for SNP1 & n1"),
])
def test_init_strats_fail_stop(args, msg):
    # the reference dies with a message and EXIT_FAILURE (share/HypJacL.cu:13-91)
    a = list(args)
    if a[0] == "S":
        a[0] = g.STRAT_SO
    code = ("import sys; sys.path.insert(0, %r); import gpujachx_amd as g; g.init_strats(*%r)"
            % (ROOT, tuple(a)))
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert msg in p.stderr
