"""Learned pair-skip on the GPU: the HIP MLP kernels against the CPU
restatement (oracle/mlp.py), and the skip machinery of the level-2 loop."""
import numpy as np
import pytest

import gpujachx_amd as g
import gpujachx_amd.learned as lr
import oracle as o
from oracle import mlp
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_form(g.FORM_PLAIN)


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


def _use_model(p, thr=0.05, max_skips=2):
    import ctypes
    assert g.lib().hj_set_skip_model(p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size,
                                     thr, max_skips) == 0
    g.set_form(g.FORM_LEARNED_SKIP)


def _plain(alg, G, nplus):
    g.set_form(g.FORM_PLAIN)
    n = G.shape[1]
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    return g.hypJacL2(alg, G, nplus)


@pytest.mark.parametrize("alg,kind", [(9, "svd"), (11, "hyp"), (3, "svd")])
def test_skip_none_is_plain(alg, kind):
    G = graded_input(512, 256, seed=21)
    nplus = 256 if kind == "svd" else 100
    rp = _plain(alg, G, nplus)
    _use_model(_constant_model(+30.0))
    rl = g.hypJacL2(alg, G, nplus)
    g.set_form(g.FORM_PLAIN)
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
    rp = _plain(9, G, 1024)
    lr.enable()
    try:
        rl = g.hypJacL2(9, G, 1024)
        sk, tot, vis = g.last_skips()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
    assert rb[-1] == 0 and sk[-1] == 0
    a, b = np.sort(rp.sigma), np.sort(rl.sigma)
    assert np.abs(a - b).max() <= 1e-13 * a.max()
