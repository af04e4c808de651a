"""Parity of the gfx950 library with the CPU oracle, through the C ABI.

The plain FP64 path is expected to match the oracle bit for bit: both run
the same IEEE-754 operations in the same order (hjoracle.h states the
order).  Every comparison below is exact (assert_array_equal; +0 == -0).
"""
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
    g.set_step_limit(0)


def run_both(alg, G, nplus, snp1="cycwor", snp0="cycwor", limit=0):
    n = G.shape[1]
    g.init_strats_builtin(snp0, 32, snp1, n // 16)
    s0, s1 = g.current_strats()
    g.set_step_limit(limit)
    try:
        rg = g.hypJacL2(alg, G, nplus)
    finally:
        g.set_step_limit(0)
    ro = o.hypJacL2(alg, G, nplus, s0, s1, max_steps=limit)
    return rg, ro


def assert_same(rg, ro, ctx=""):
    assert rg.ret == 0 and ro["ret"] == 0, ctx
    assert rg.sweeps == ro["sweeps"], ctx
    assert (rg.rot_s, rg.rot_b) == (ro["rot_s"], ro["rot_b"]), ctx
    np.testing.assert_array_equal(rg.D, ro["D"], err_msg=ctx)
    np.testing.assert_array_equal(rg.G, ro["G"], err_msg=ctx)
    if ro["V"] is not None:
        np.testing.assert_array_equal(rg.V, ro["V"], err_msg=ctx)


def test_scalar_kernels_bitwise():
    rng = np.random.default_rng(11)
    n = 4096
    x = np.concatenate([rng.uniform(0.5, 4.0, n), 10.0 ** rng.uniform(-300, 300, n),
                        1.0 + rng.uniform(0, 1, n) ** 8, np.array([1.0, 2.0, 4.0, 0.25, 5e-324, 1e308])])
    # plus a million mantissas over [1,4) and values next to powers of 4
    m = rng.uniform(1.0, 4.0, 1 << 20)
    edge = np.array([np.nextafter(4.0 ** k, 0.0) for k in range(-3, 4)] +
                    [np.nextafter(4.0 ** k, np.inf) for k in range(-3, 4)])
    # subnormals of every exponent, the extremes, special values
    sub = rng.uniform(1.0, 2.0, 52 * 8) * np.repeat(2.0 ** -np.arange(1023, 1075, dtype=float), 8)
    spec = np.array([0.0, -0.0, -1.0, -5e-324, np.inf, -np.inf, np.nan, 5e-324, 1e-310,
                     np.nextafter(2.0 ** -1022, 0.0), 2.0 ** -1022, np.finfo(float).max])
    x = np.concatenate([x, m, m * 2.0 ** -700, edge, sub, spec])
    c, _, _ = g.dbg_scalar(0, x)
    np.testing.assert_array_equal(c, o.rsqrt_n(x))
    # rotations across all branches: |ct2| tiny, moderate, huge, around 1
    apq = rng.standard_normal(3 * n) * 10.0 ** rng.uniform(-10, 2, 3 * n)
    dp = 10.0 ** rng.uniform(-3, 3, 3 * n)
    dq = dp * (1.0 + 10.0 ** rng.uniform(-12, 1, 3 * n) * rng.choice([-1, 1], 3 * n))
    for kind, fn in ((1, o.rot_t), (2, o.rot_h)):
        cg, tg, bg = g.dbg_scalar(kind, apq, dp, dq)
        for i in range(apq.size):
            co, to, bo = fn(apq[i], dp[i], dq[i])
            assert (cg[i], tg[i], bg[i]) == (co, to, bo), (kind, i, apq[i], dp[i], dq[i])


@pytest.mark.parametrize("alg", list(range(16)))
@pytest.mark.parametrize("kind", ["svd", "neg", "hyp"])
def test_one_step_every_alg(alg, kind):
    m, n = 256, 128
    G = graded_input(m, n, seed=100 + alg, lo=-2, hi=2)
    nplus = {"svd": n, "neg": 0, "hyp": 40}[kind]
    rg, ro = run_both(alg, G, nplus, limit=1)
    assert_same(rg, ro, "alg %d %s" % (alg, kind))


@pytest.mark.parametrize("alg", [9, 11])
def test_c1_svd_full(alg):
    # the plain-method SVD workload: M=N=256, s_i = 10^(-6(i-1)/255), cycwor/cycwor
    n = 256
    s = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    G = svd_input(n, n, s, seed=1)
    rg, ro = run_both(alg, G, n)
    assert_same(rg, ro, "C1 alg %d" % alg)
    sig = np.sort(rg.sigma)[::-1]
    assert np.max(np.abs(sig - s) / s) < 1e-9


@pytest.mark.parametrize("alg", [0, 1, 2, 3, 8, 10])
def test_small_full_other_algs(alg):
    G = svd_input(256, 128, 10.0 ** (-4.0 * np.arange(128) / 127), seed=7)
    rg, ro = run_both(alg, G, 128)
    assert_same(rg, ro, "alg %d" % alg)


@pytest.mark.parametrize("nplus", [0, 16, 37, 64, 128])
def test_hsvd_full(nplus):
    m, n = 512, 128
    G = graded_input(m, n, seed=2 + nplus)
    rg, ro = run_both(11, G, nplus)
    assert_same(rg, ro, "hsvd nplus %d" % nplus)
    J = np.diag(np.r_[np.ones(nplus), -np.ones(n - nplus)])
    V = rg.V
    assert np.abs(V.T @ J @ V - J).max() < 1e-10 * max(1.0, np.abs(V).max() ** 2)


@pytest.mark.parametrize("snp", ["BrentL", "rowcyc", "colcyc", "cycloc", "mmstep"])
def test_strategies_full(snp):
    G = svd_input(256, 256, 10.0 ** (-3.0 * np.arange(256) / 255), seed=3)
    rg, ro = run_both(9, G, 256, snp1=snp, snp0=snp)
    assert_same(rg, ro, snp)


@pytest.mark.parametrize("alg", [9, 1, 8])
@pytest.mark.parametrize("descending", [True, False])
def test_identity_updates_left_out(alg, descending):
    # mutually orthogonal columns (disjoint rows): every Gram is diagonal, no
    # pivot pair rotates; in the sorted variants (ALG 9, 1) unsorted norms
    # make the sort rule swap columns (an update by a permutation), sorted
    # ones leave W the identity and the update out.  GPU and oracle agree on
    # every bit, and a run without swaps leaves every column in its place.
    m, n = 256, 64
    d = 10.0 ** -np.linspace(0.0, 6.0, n)
    if not descending:
        d = d[::-1].copy()
    rows = np.random.default_rng(5).permutation(m)[:n]
    G = np.zeros((m, n))
    G[rows, np.arange(n)] = d
    rg, ro = run_both(alg, G, n)
    assert_same(rg, ro, "alg %d descending %s" % (alg, descending))
    assert rg.rot_s == 0
    # the output columns are normalised; without swaps each keeps its row
    same_pattern = np.array_equal(rg.G != 0.0, G != 0.0)
    assert same_pattern == (descending or not (alg & 1))
