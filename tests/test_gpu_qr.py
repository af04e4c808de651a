"""The QR factorization path and Drmac's sorted level-0 on the GPU against
the oracle, bit for bit (every element of U, V, D, the counters, the sweep
count), through the C ABI."""
import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_factorization(g.FACT_CHOLESKY)
    o.set_factorization(o.FACT_CHOLESKY)


def run_both(fact, alg, G, nplus, snp1="cycwor"):
    n = G.shape[1]
    g.init_strats_builtin("cycwor", 32, snp1, n // 16)
    s0, s1 = g.current_strats()
    g.set_factorization(fact)
    o.set_factorization(fact)
    try:
        rg = g.hypJacL2(alg, G, nplus)
        ro = o.hypJacL2(alg, G, nplus, s0, s1, nthreads=8)
    finally:
        g.set_factorization(g.FACT_CHOLESKY)
        o.set_factorization(o.FACT_CHOLESKY)
    return rg, ro


def assert_same(rg, ro, ctx=""):
    assert rg.ret == 0 and ro["ret"] == 0, ctx
    assert rg.sweeps == ro["sweeps"], ctx
    assert (rg.rot_s, rg.rot_b) == (ro["rot_s"], ro["rot_b"]), ctx
    np.testing.assert_array_equal(rg.D, ro["D"], err_msg=ctx)
    np.testing.assert_array_equal(rg.G, ro["G"], err_msg=ctx)
    if ro["V"] is not None:
        np.testing.assert_array_equal(rg.V, ro["V"], err_msg=ctx)


@pytest.mark.parametrize("fact", [g.FACT_QR, g.FACT_QR_DRMAC])
@pytest.mark.parametrize("alg", [9, 11, 8, 1, 0, 3])
@pytest.mark.parametrize("kind", ["svd", "hyp", "neg"])
def test_qr_paths_bitwise(fact, alg, kind):
    m, n = 256, 128
    if kind == "hyp":
        G, nplus = graded_input(m, n, seed=31 + alg), 40
    else:
        s = 10.0 ** (-10.0 * np.random.default_rng(alg).uniform(size=n))
        G, nplus = svd_input(m, n, s, seed=30 + alg), (n if kind == "svd" else 0)
    rg, ro = run_both(fact, alg, G, nplus)
    assert_same(rg, ro, "fact %d alg %d %s" % (fact, alg, kind))


@pytest.mark.parametrize("fact", [g.FACT_QR, g.FACT_QR_DRMAC])
def test_qr_paths_larger_and_graded(fact):
    # 1024 rows (32 panels per pair), column scales 10^U(-60, 60)
    G = graded_input(1024, 256, seed=5, lo=-60.0, hi=60.0)
    rg, ro = run_both(fact, 9, G, 256)
    assert_same(rg, ro, "graded")
    rg, ro = run_both(fact, 11, G, 100, snp1="rowcyc")
    assert_same(rg, ro, "graded hsvd rowcyc")


def test_qr_path_ignores_the_forms():
    # the learned / split forms act on the Gram of the Cholesky path: with QR
    # the run is the plain QR run
    G = graded_input(256, 128, seed=8)
    g.set_form(g.FORM_SPLIT_GRAM)
    try:
        rg, ro = run_both(g.FACT_QR_DRMAC, 9, G, 128)
    finally:
        g.set_form(g.FORM_PLAIN)
    assert_same(rg, ro, "split form set")
