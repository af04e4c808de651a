"""The wide-block form (blocks of 32 columns, 64-column pairs) on the GPU:
bit for bit against its oracle (every bit of G, V, D and the counters)
through the C ABI, over ALGs, J kinds, strategy families and sizes."""
import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    old = g.lib().hj_set_block_width(32)
    yield
    g.lib().hj_set_block_width(old)


def _both(alg, G, nplus, fam="cycwor", ldhG=None):
    n = G.shape[1]
    g.init_strats_builtin(fam, 32, fam, n // 16)
    rg = g.hypJacL2(alg, G, nplus, ldhG=ldhG)
    s0 = g.strat_generate(fam, 64).astype(np.uint8)
    s1 = g.strat_generate(fam, n // 32)
    ro = o.hypJacL2_w(alg, G, nplus, 32, s0, s1, max_sweeps=30)
    return rg, ro


def assert_same(rg, ro, ctx=""):
    assert rg.ret == 0 and ro["ret"] == 0, ctx
    assert rg.sweeps == ro["sweeps"], ctx
    assert (rg.rot_s, rg.rot_b) == (ro["rot_s"], ro["rot_b"]), ctx
    np.testing.assert_array_equal(rg.D.view(np.uint64), ro["D"].view(np.uint64), err_msg=ctx)
    np.testing.assert_array_equal(rg.G.view(np.uint64), ro["G"].view(np.uint64), err_msg=ctx)
    if ro["V"] is not None:
        np.testing.assert_array_equal(rg.V.view(np.uint64), ro["V"].view(np.uint64), err_msg=ctx)


@pytest.mark.parametrize("alg", [0, 1, 2, 3, 8, 9, 10, 11])
@pytest.mark.parametrize("nplus", [128, 0, 40, 64, 100])
def test_wide_equals_oracle(alg, nplus):
    G = graded_input(256, 128, seed=100 + nplus, lo=-3, hi=3)
    print("case: wide alg %d nplus %d" % (alg, nplus), flush=True)
    rg, ro = _both(alg, G, nplus)
    assert_same(rg, ro, "alg %d nplus %d" % (alg, nplus))
    assert rg.rot_b == ro["rot_b"] and ro["rot_b_trace"][-1] == 0


@pytest.mark.parametrize("fam", ["rowcyc", "colcyc", "mmstep"])
def test_wide_strategy_families(fam):
    G = graded_input(512, 256, seed=7, lo=-4, hi=4)
    print("case: wide family %s" % fam, flush=True)
    rg, ro = _both(11, G, 96, fam=fam)
    assert_same(rg, ro, fam)


@pytest.mark.parametrize("alg,nplus", [(9, 512), (11, 512), (11, 200), (3, 300)])
def test_wide_larger(alg, nplus):
    s = 10.0 ** (-10.0 * np.random.default_rng(1).uniform(size=512))
    G = svd_input(1024, 512, s, seed=3) if nplus == 512 else graded_input(1024, 512, seed=5)
    print("case: wide 1024x512 alg %d nplus %d" % (alg, nplus), flush=True)
    rg, ro = _both(alg, G, nplus)
    assert_same(rg, ro, "1024x512 alg %d nplus %d" % (alg, nplus))


@pytest.mark.parametrize("kind", ["svd", "hyp"])
def test_wide_learned_skip(kind):
    # the learned pair-skip on 32-column blocks (the packaged model, the
    # run-level floors off): visits skipped, the run ends on an unskipped
    # sweep without a big rotation, the spectrum per value as the plain run's
    import os
    import gpujachx_amd.learned as lr
    from conftest import assert_per_value
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 1024
    s = 10.0 ** (-10.0 * np.random.default_rng(3).uniform(size=n))
    G = svd_input(n, n, s, seed=31)
    nplus = n if kind == "svd" else n // 2
    old = g.set_max_sweeps(80)
    try:
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        rp = g.hypJacL2(9, G, nplus)
        g.init_strats_builtin("cycwor", 32, "rowcyc", n // 16)
        rr = g.hypJacL2(9, G, nplus)
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        lr.enable(form=g.FORM_LEARNED_SKIP)
        g.lib().hj_set_skip_min_visits(0)
        g.lib().hj_set_skip_min_propose(0.0)
        print("case: wide learned %s" % kind, flush=True)
        rl = g.hypJacL2(9, G, nplus)
        sk, skipped, visits = g.last_skips()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old)
        g.lib().hj_set_skip_min_visits(32000)
        g.lib().hj_set_skip_min_propose(0.1)
    assert rp.ret == 0 and rl.ret == 0
    print("plain %d sweeps, learned %d sweeps, %d of %d visits skipped"
          % (rp.sweeps, rl.sweeps, skipped, skipped + visits))
    assert skipped > 0
    assert rb[-1] == 0 and sk[-1] == 0
    assert_per_value(rl.D, rp.D, rr.D, "wide learned %s" % kind)


def test_wide_ignores_split_and_qr():
    # at width 32 the split-Gram bit of the form and the QR factorization
    # are ignored: the plain FP64 run's bits
    G = graded_input(512, 256, seed=21)
    g.init_strats_builtin("cycwor", 32, "cycwor", 256 // 16)
    rp = g.hypJacL2(11, G, 100)
    g.set_form(g.FORM_SPLIT_GRAM)
    try:
        rs = g.hypJacL2(11, G, 100)
    finally:
        g.set_form(g.FORM_PLAIN)
    old = g.set_factorization(g.FACT_QR)
    try:
        rq = g.hypJacL2(11, G, 100)
    finally:
        g.set_factorization(old)
    for r in (rs, rq):
        assert r.ret == 0 and (r.sweeps, r.rot_s, r.rot_b) == (rp.sweeps, rp.rot_s, rp.rot_b)
        for a, b in ((r.G, rp.G), (r.V, rp.V), (r.D, rp.D)):
            np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


def test_wide_padded_leading_dimension():
    G = graded_input(320, 128, seed=9)
    print("case: wide ldhG 400", flush=True)
    rg, ro = _both(11, G, 70, ldhG=400)
    assert_same(rg, ro, "ldhG 400")


def test_wide_order_without_table():
    # 192 columns: level 1 of order 6 at width 32; mmstep has it, BrentL
    # loaded at order 12 (width 16) too; a family without a table of the
    # order this width needs gives -3 (init_strats itself dies on such orders)
    G = graded_input(256, 192, seed=10)
    print("case: wide BrentL 192", flush=True)
    rg, ro = _both(9, G, 192, fam="BrentL")
    assert_same(rg, ro, "BrentL 192")
    g.init_strats_builtin("cycwor", 32, "cycwor", 128 // 16)
    assert g.hypJacL2(9, graded_input(256, 192, seed=11), 192).ret == -3


def test_wide_argument_codes_and_width_switch():
    G = graded_input(256, 96, seed=2)
    g.init_strats_builtin("BrentL", 32, "BrentL", 96 // 16)
    assert g.hypJacL2(9, G, 96).ret == -3           # not a multiple of 64 columns
    assert g.lib().hj_set_block_width(7) == 32      # ignored
    old = g.lib().hj_set_block_width(16)
    try:
        r = g.hypJacL2(9, G, 96)                    # the reference's width takes it
        s0 = g.strat_generate("BrentL", 32).astype(np.uint8)
        s1 = g.strat_generate("BrentL", 96 // 16)
        ro = o.hypJacL2(9, G, 96, s0, s1)
        assert_same(r, ro, "width 16")
    finally:
        g.lib().hj_set_block_width(old)


def test_wide_spectrum_against_reference_width():
    # the two widths are different FP64 methods: the same spectrum to FP64
    # accuracy on a graded matrix of N = 1024
    from conftest import assert_per_value
    G = graded_input(1024, 1024, seed=4)
    n = 1024
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    rw = g.hypJacL2(11, G, 512)
    old = g.lib().hj_set_block_width(16)
    try:
        r16 = g.hypJacL2(11, G, 512)
        g.init_strats_builtin("cycwor", 32, "rowcyc", n // 16)
        rr = g.hypJacL2(11, G, 512)
    finally:
        g.lib().hj_set_block_width(old)
    assert rw.ret == 0 and r16.ret == 0
    print("wide %d sweeps, width 16 %d sweeps" % (rw.sweeps, r16.sweeps))
    assert_per_value(rw.D, r16.D, rr.D, "wide vs width 16")
