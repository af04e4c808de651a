"""Learned pair-skip on the GPU: the HIP MLP kernels against the CPU
restatement (oracle/mlp.py), and the skip machinery of the level-2 loop."""
import numpy as np
import pytest

import gpujachx_amd as g
import gpujachx_amd.learned as lr
import oracle as o
from oracle import mlp
from conftest import assert_per_value, graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_form(g.FORM_PLAIN)
    g.lib().hj_set_skip_min_visits(32000)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, mlp.F)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.float32)
    p = (rng.standard_normal(mlp.NPARAM) * 0.2).astype(np.float32)
    return X, y, p


@pytest.mark.parametrize("n", [1, 257, 4096])
def test_mlp_grad_matches_oracle(n):
    X, y, p = _data(n, n)
    gg, loss = lr.grad(X, y, p, wpos=3.0)
    lo, go = mlp.loss_and_grad(X, y, mlp.bf16_round(p), 3.0)
    # FP32 accumulation over the batch vs FP64
    tol = 2e-5 * np.abs(go).max() + 1e-7
    assert np.abs(gg - go).max() <= tol, np.abs(gg - go).max() / np.abs(go).max()
    assert abs(loss - lo) <= 1e-5 * max(1.0, lo)


def test_mlp_predict_matches_oracle():
    X, _, p = _data(5000, 9)
    pg = lr.predict(X, p)
    po = mlp.predict(X, mlp.bf16_round(p))
    assert np.abs(pg - po).max() < 1e-5


def test_features_match_oracle():
    rng = np.random.default_rng(4)
    k = 3000
    rec = np.zeros(k, dtype=mlp.REC_DTYPE)
    rec["rot_b"] = rng.integers(0, 600, k)
    rec["rot_s"] = rec["rot_b"] + rng.integers(0, 50, k)
    rec["maxcos"] = (10.0 ** rng.uniform(-20, 0, k)).astype(np.float32)
    rec["l2ni"] = rng.uniform(-60, 60, k).astype(np.float32)
    rec["l2nj"] = rng.uniform(-60, 60, k).astype(np.float32)
    rec["sweep"] = rng.integers(1, 9, k)
    rec["skips"] = rng.integers(0, 3, k)
    s1 = rng.integers(0, 512, (k, 2)).astype(np.uint16)
    Xg = lr.features(rec, s1, 1000, 512, 10, 3e-4)
    Xo = mlp.features(rec, s1, 1000, 512, 10, 3e-4)
    assert np.abs(Xg - Xo).max() < 2e-6


def test_training_learns_a_rule():
    # label = big rotation iff feature 0 > 0.2 or feature 2 > -0.25
    rng = np.random.default_rng(5)
    n = 60000
    X = rng.uniform(-1, 1, (n, mlp.F)).astype(np.float32)
    y = ((X[:, 0] > 0.2) | (X[:, 2] > -0.25)).astype(np.float32)
    p, losses = lr.train(X, y, epochs=4, batch=1024, lr=3e-3, wpos=1.0)
    assert losses[-1] < 0.5 * losses[0]
    acc = np.mean((lr.predict(X, p) > 0.5) == (y > 0.5))
    assert acc > 0.97, acc


def _constant_model(logit):
    p = np.zeros(mlp.NPARAM, dtype=np.float32)
    p[mlp.OFF_B3] = logit
    return p


def _use_model(p, thr=0.05, max_skips=2, min_frac=0.0, min_propose=0.0):
    import ctypes
    assert g.lib().hj_set_skip_model(p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size,
                                     thr, max_skips) == 0
    # skipping allowed in every sweep and run (the policy is tested apart)
    g.lib().hj_set_skip_min_visits(0)
    g.lib().hj_set_skip_min_frac(min_frac)
    g.lib().hj_set_skip_min_propose(min_propose)
    g.set_form(g.FORM_LEARNED_SKIP)


def _plain(alg, G, nplus, snp1="cycwor"):
    g.set_form(g.FORM_PLAIN)
    n = G.shape[1]
    g.init_strats_builtin("cycwor", 32, snp1, n // 16)
    r = g.hypJacL2(alg, G, nplus)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    return r


@pytest.mark.parametrize("alg,kind", [(9, "svd"), (11, "hyp"), (3, "svd")])
def test_skip_none_is_plain(alg, kind):
    G = graded_input(512, 256, seed=21)
    nplus = 256 if kind == "svd" else 100
    rp = _plain(alg, G, nplus)
    _use_model(_constant_model(+30.0))
    rl = g.hypJacL2(alg, G, nplus)
    g.set_form(g.FORM_PLAIN)
    g.lib().hj_set_skip_min_frac(0.01)
    g.lib().hj_set_skip_min_propose(0.1)
    assert rl.sweeps == rp.sweeps and rl.rot_b == rp.rot_b
    np.testing.assert_array_equal(rl.G, rp.G)
    np.testing.assert_array_equal(rl.D, rp.D)
    sk, tot, _ = g.last_skips()
    assert tot == 0


@pytest.mark.parametrize("alg,kind", [(9, "svd"), (11, "hyp")])
def test_skip_all_is_plain_with_idle_sweeps(alg, kind):
    # every visit the model may skip is skipped: a skipped sweep changes
    # nothing, so the executed sweeps are the plain run's, bit for bit, and
    # the run still ends on an unskipped sweep without big rotations
    G = graded_input(512, 256, seed=22)
    nplus = 256 if kind == "svd" else 100
    rp = _plain(alg, G, nplus)
    s0, s1 = g.current_strats()
    ro = o.hypJacL2(alg, G, nplus, s0, s1)
    np.testing.assert_array_equal(rp.G, ro["G"])
    old = g.set_max_sweeps(200)
    try:
        _use_model(_constant_model(-30.0), max_skips=1)
        rl = g.hypJacL2(alg, G, nplus)
        sk, tot, vis = g.last_skips()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old)
        g.lib().hj_set_skip_min_frac(0.01)
        g.lib().hj_set_skip_min_propose(0.1)
    np.testing.assert_array_equal(rl.G, rp.G)
    np.testing.assert_array_equal(rl.D, rp.D)
    if rl.V is not None:
        np.testing.assert_array_equal(rl.V, rp.V)
    assert rl.rot_s == rp.rot_s and rl.rot_b == rp.rot_b
    assert sk[-1] == 0 and sk[0] == 0 and tot > 0
    assert rl.sweeps == 2 * rp.sweeps - 1
    keys = (256 // 32) * (256 // 16 - 1)
    assert vis == rp.sweeps * keys


def test_learned_model_file_converges():
    import os
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    G = svd_input(1024, 1024, 10.0 ** (-10.0 * np.random.default_rng(3).uniform(size=1024)), seed=31)
    rr = _plain(9, G, 1024, snp1="rowcyc")
    rp = _plain(9, G, 1024)
    lr.enable()
    g.lib().hj_set_skip_min_visits(0)
    try:
        rl = g.hypJacL2(9, G, 1024)
        sk, tot, vis = g.last_skips()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
    assert rb[-1] == 0 and sk[-1] == 0
    a, b = np.sort(rp.sigma), np.sort(rl.sigma)
    assert np.abs(a - b).max() <= 1e-13 * a.max()
    assert_per_value(rl.D, rp.D, rr.D, "packaged model N=1024")


def test_skip_policy_gates():
    # with min_propose above 1 no sweep may skip: the run is the plain run;
    # with min_frac above 1 neither
    G = graded_input(512, 256, seed=23)
    rp = _plain(9, G, 256)
    for mf, mp in ((0.0, 1.5), (1.5, 0.0)):
        _use_model(_constant_model(-30.0), min_frac=mf, min_propose=mp)
        try:
            rl = g.hypJacL2(9, G, 256)
            _, tot, _ = g.last_skips()
        finally:
            g.set_form(g.FORM_PLAIN)
            g.lib().hj_set_skip_min_frac(0.01)
            g.lib().hj_set_skip_min_propose(0.1)
        assert tot == 0
        np.testing.assert_array_equal(rl.G, rp.G)
        assert rl.sweeps == rp.sweeps


def test_small_runs_do_not_skip():
    # below the visits-per-sweep floor the learned form runs the plain method
    G = graded_input(512, 256, seed=24)
    rp = _plain(9, G, 256)
    _use_model(_constant_model(-30.0))
    g.lib().hj_set_skip_min_visits(32000)
    try:
        rl = g.hypJacL2(9, G, 256)
        _, tot, _ = g.last_skips()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.lib().hj_set_skip_min_frac(0.01)
        g.lib().hj_set_skip_min_propose(0.1)
    assert tot == 0 and rl.sweeps == rp.sweeps
    np.testing.assert_array_equal(rl.G, rp.G)


def test_learned_gain_at_4096():
    # what the trained network gains: at N = 4096 (above the visits floor)
    # the packaged model executes clearly fewer block-pair visits than the
    # plain method, ends on an unskipped converged sweep, and gives the same
    # sigma; both wall times are printed
    import os
    import time
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 4096
    s = 10.0 ** (-12.0 * np.random.default_rng(8).uniform(size=n))
    G = svd_input(n, n, s, seed=33)
    old = g.set_max_sweeps(80)
    try:
        rr = _plain(9, G, n, snp1="rowcyc")
        t0 = time.perf_counter()
        rp = _plain(9, G, n)
        tp = time.perf_counter() - t0
        _, _, vp = g.last_skips()
        lr.enable()
        t0 = time.perf_counter()
        rl = g.hypJacL2(9, G, n)
        tl = time.perf_counter() - t0
        sk, tot, vl = g.last_skips()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old)
    print("N=4096: plain %.3f s %d sweeps %d visits; learned %.3f s %d sweeps %d visits (%d skipped)"
          % (tp, rp.sweeps, vp, tl, rl.sweeps, vl, tot))
    assert rb[-1] == 0 and sk[-1] == 0
    assert vl <= 0.9 * vp, (vl, vp)
    a, b = np.sort(rp.sigma), np.sort(rl.sigma)
    assert np.abs(a - b).max() <= 1e-13 * a.max()
    print("  per value vs plain %.3g, FP64 spread %.3g" % assert_per_value(rl.D, rp.D, rr.D,
                                                                           "N=4096"))
