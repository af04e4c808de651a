"""The learned pair-skip configuration's matrix class (G = U diag(s) V^T
diag(10^u): Haar U, V, s log-uniform in [1e-10, 1], column grading
u_j ~ U(-4, 4)), SVD and hyperbolic SVD with NPLUS = N/2, at orders the
oracle finishes in seconds:

* the plain FP64 path equals the oracle bit for bit on it (N = 512);
* the packaged model (learned skip + split Gram, the visits-per-sweep floor
  off) skips visits, ends on an unskipped all-FP64 sweep without a big
  rotation, and its values agree with the plain path's per value, within
  ten times the spread of two FP64 runs that differ only in the strategy
  (cycwor against rowcyc), the large ones to 1e-11 -- a wrong small value
  fails even where the normwise measure cannot see it (N = 1024).
"""
import os

import numpy as np
import pytest

import gpujachx_amd as g
import gpujachx_amd.learned as lr
import oracle as o
from gpujachx_amd.accuracy import per_sigma, rel_diffs
from conftest import haar

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_form(g.FORM_PLAIN)
    g.lib().hj_set_skip_min_visits(32000)
    g.lib().hj_set_skip_min_propose(0.1)


def c5_input(n, seed):
    rng = np.random.default_rng(seed)
    U, V = haar(n, rng), haar(n, rng)
    s = 10.0 ** (-10.0 * rng.uniform(size=n))
    return (U * s) @ V.T * 10.0 ** rng.uniform(-4.0, 4.0, n)


def _plain(G, nplus, snp1="cycwor"):
    g.set_form(g.FORM_PLAIN)
    g.init_strats_builtin("cycwor", 32, snp1, G.shape[1] // 16)
    return g.hypJacL2(9, G, nplus)


@pytest.mark.parametrize("kind", ["svd", "hsvd"])
def test_plain_equals_oracle(kind):
    n = 512
    G = c5_input(n, seed=51)
    nplus = n if kind == "svd" else n // 2
    r = _plain(G, nplus)
    s0, s1 = g.current_strats()
    ro = o.hypJacL2(9, G, nplus, s0, s1, nthreads=8)
    assert r.ret == 0 and ro["ret"] == 0
    assert r.sweeps == ro["sweeps"] and (r.rot_s, r.rot_b) == (ro["rot_s"], ro["rot_b"])
    np.testing.assert_array_equal(r.D, ro["D"])
    np.testing.assert_array_equal(r.G, ro["G"])


@pytest.mark.parametrize("kind", ["svd", "hsvd"])
def test_packaged_model(kind):
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 1024
    G = c5_input(n, seed=52)
    nplus = n if kind == "svd" else n // 2
    old = g.set_max_sweeps(80)
    try:
        rp = _plain(G, nplus)
        rr = _plain(G, nplus, snp1="rowcyc")
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        lr.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
        # the floor off: skipping in a run of this size, on any proposal
        g.lib().hj_set_skip_min_visits(0)
        g.lib().hj_set_skip_min_propose(0.0)
        rl = g.hypJacL2(9, G, nplus)
        sk, skipped, visits = g.last_skips()
        sp, split = g.last_splits()
        _, rb, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old)
        g.lib().hj_set_skip_min_visits(32000)
        g.lib().hj_set_skip_min_propose(0.1)
    assert rp.ret == 0 and rr.ret == 0 and rl.ret == 0
    fast = per_sigma(rl.D, rp.D)
    spread = per_sigma(rr.D, rp.D)
    print("%s N=%d: plain %d sweeps, rowcyc %d, learned+split %d sweeps, %d of %d visits "
          "skipped, %d split Grams; per value vs plain %s; FP64 spread %s"
          % (kind, n, rp.sweeps, rr.sweeps, rl.sweeps, skipped, skipped + visits, split, fast,
             spread))
    # the forms acted, and the run ended the reference's way
    assert skipped > 0 and split > 0
    assert rb[-1] == 0 and sk[-1] == 0 and sp[-1] == 0
    # the same inertia, and per value within ten times the FP64 spread
    assert "inertia_mismatch" not in fast
    rel_f, ref = rel_diffs(rl.D, rp.D)
    rel_r, _ = rel_diffs(rr.D, rp.D)
    assert rel_f.max() <= max(10.0 * rel_r.max(), 1e-12), (fast, spread)
    # the large values (at least a tenth of the largest), well conditioned
    # normwise, agree to 1e-11
    big = ref >= 0.1 * ref.max()
    assert np.all(rel_f[big] <= 1e-11), rel_f[big].max()
    assert fast["normwise"] <= 1e-13
