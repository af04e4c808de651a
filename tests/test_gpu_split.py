"""Split-precision Gram (BF16 MFMA, three-way splits, FP32 accumulation,
FP64 recombination) and the split form of the method, on the GPU."""
import ctypes

import numpy as np
import pytest

import gpujachx_amd as g
from conftest import assert_per_value, graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_form(g.FORM_PLAIN)
    g.set_split_tau(1e-3)


@pytest.mark.parametrize("nrow", [64, 1024, 4096])
def test_split_gram_exact_on_small_integers(nrow):
    # integer data |x| <= 15: every bf16 piece product and every 64-row FP32
    # sum is exact, so the split Gram equals the FP64 Gram exactly; this pins
    # the lane layouts, the row coverage and the scaling of the split kernel
    rng = np.random.default_rng(nrow)
    G = rng.integers(-15, 16, (nrow, 32)).astype(np.float64)
    G[:, 5] *= 2.0 ** 40      # power-of-two column scales stay exact
    G[:, 20] *= 2.0 ** -40
    exact = G.T @ G
    A64 = g.dbg_gram(G, split=False)
    As = g.dbg_gram(G, split=True)
    for A in (A64, As):
        np.testing.assert_array_equal(np.tril(A), np.tril(exact))


@pytest.mark.parametrize("nrow,lo,hi", [(1024, -4, 4), (8192, -12, 0), (2048, -150, 150)])
def test_split_gram_relative_error(nrow, lo, hi):
    rng = np.random.default_rng(7)
    G = rng.standard_normal((nrow, 32)) * 10.0 ** rng.uniform(lo, hi, 32)
    A64 = g.dbg_gram(G, split=False)
    As = g.dbg_gram(G, split=True)
    d = np.sqrt(np.diag(A64))
    rel = np.abs(As - A64) / np.outer(d, d)
    # three-way split: about FP32 accuracy per 64-row group, FP64 across
    assert rel.max() < 2e-6, rel.max()
    ref = G.T @ G
    assert (np.abs(A64 - ref) / np.outer(d, d)).max() < 1e-13


def _run(form, alg, G, nplus, tau=1e-3, snp1="cycwor"):
    n = G.shape[1]
    g.init_strats_builtin("cycwor", 32, snp1, n // 16)
    g.set_form(form)
    g.set_split_tau(tau)
    try:
        r = g.hypJacL2(alg, G, nplus)
        sp, tot = g.last_splits()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
    return r, sp, tot, rb


@pytest.mark.parametrize("alg,kind", [(9, "svd"), (11, "svd"), (11, "hyp"), (3, "svd")])
def test_split_form_matches_plain(alg, kind):
    n = 256
    if kind == "svd":
        G = svd_input(512, n, 10.0 ** (-10.0 * np.random.default_rng(1).uniform(size=n)), seed=41)
        nplus = n
    else:
        G = graded_input(512, n, seed=42)
        nplus = 100
    rp, _, _, _ = _run(g.FORM_PLAIN, alg, G, nplus)
    rr, _, _, _ = _run(g.FORM_PLAIN, alg, G, nplus, snp1="rowcyc")
    rs, sp, tot, rb = _run(g.FORM_SPLIT_GRAM, alg, G, nplus)
    assert tot > 0 and sp[0] > 0
    assert sp[-1] == 0 and rb[-1] == 0          # ends on an all-FP64 sweep
    a, b = np.sort(rp.sigma), np.sort(rs.sigma)
    assert np.abs(a - b).max() <= 1e-13 * a.max()
    assert_per_value(rs.D, rp.D, rr.D, "split alg %d %s" % (alg, kind))
    if rs.V is not None:
        J = np.diag(np.r_[np.ones(nplus), -np.ones(n - nplus)])
        assert np.abs(rs.V.T @ J @ rs.V - J).max() < 1e-11 * max(1.0, np.abs(rs.V).max() ** 2)
        # residual G V = U Sigma
        res = G @ rs.V - rs.G * rs.sigma
        assert np.abs(res).max() < 1e-12 * np.abs(G).max() * n


def test_fused_split_grams():
    # the fused block update + next-step Gram forms the split Gram of every
    # split-planned pair: in the first sweep (every visit planned split) all
    # visits then hold split Grams; with it off only the visits after a
    # standalone Gram launch do
    n = 512
    G = svd_input(512, n, 10.0 ** (-10.0 * np.random.default_rng(4).uniform(size=n)), seed=44)
    rp, _, _, _ = _run(g.FORM_PLAIN, 9, G, n)
    rr, _, _, _ = _run(g.FORM_PLAIN, 9, G, n, snp1="rowcyc")
    r0, sp0, tot0, rb0 = _run(g.FORM_SPLIT_GRAM, 9, G, n)
    old = g.lib().hj_set_fused_split(1)
    try:
        r1, sp1, tot1, rb1 = _run(g.FORM_SPLIT_GRAM, 9, G, n)
    finally:
        g.lib().hj_set_fused_split(old)
    assert old == 0
    visits = (n // 32) * (n // 16 - 1)
    assert sp1[0] == visits, (sp1[0], visits)
    assert 0 < sp0[0] < visits and tot0 < tot1
    for r, sp, rb in ((r0, sp0, rb0), (r1, sp1, rb1)):
        assert r.ret == 0 and rb[-1] == 0 and sp[-1] == 0
        assert_per_value(r.D, rp.D, rr.D, "split form, fused split Grams %s" % (r is r1))


def test_split_form_nearly_dependent_columns():
    # two columns at an angle ~1e-6: their split Gram may lose positive
    # definiteness; such visits must fall back to FP64 and the result match
    rng = np.random.default_rng(5)
    G = rng.standard_normal((256, 64))
    G[:, 17] = G[:, 3] + 1e-6 * rng.standard_normal(256)
    rp, _, _, _ = _run(g.FORM_PLAIN, 11, G, 64)
    rs, sp, tot, rb = _run(g.FORM_SPLIT_GRAM, 11, G, 64)
    assert rb[-1] == 0 and sp[-1] == 0
    a, b = np.sort(rp.sigma), np.sort(rs.sigma)
    assert np.abs(a - b).max() <= 1e-12 * a.max()


def test_learned_plus_split_matches_plain():
    import os
    import gpujachx_amd.learned as lr
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    G = svd_input(1024, 1024, 10.0 ** (-12.0 * np.random.default_rng(2).uniform(size=1024)), seed=43)
    old = g.set_max_sweeps(80)
    try:
        rp, _, _, rbp = _run(g.FORM_PLAIN, 9, G, 1024)
        rr, _, _, _ = _run(g.FORM_PLAIN, 9, G, 1024, snp1="rowcyc")
        g.init_strats_builtin("cycwor", 32, "cycwor", 64)
        lr.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
        g.lib().hj_set_skip_min_visits(0)
        rl = g.hypJacL2(9, G, 1024)
        sp, _ = g.last_splits()
        sk, _, _ = g.last_skips()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old)
        g.lib().hj_set_skip_min_visits(32000)
    assert rbp[-1] == 0
    assert rb[-1] == 0 and sp[-1] == 0 and sk[-1] == 0
    a, b = np.sort(rp.sigma), np.sort(rl.sigma)
    assert np.abs(a - b).max() <= 1e-13 * a.max()
    assert_per_value(rl.D, rp.D, rr.D, "learned+split N=1024")
