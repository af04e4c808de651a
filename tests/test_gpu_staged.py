"""Staged sweeps (each step transition as two launches, half the closed quads'
fused block updates beside the pair visits the previous launch prepared)
against the step-by-step launches and the oracle: every element of U, V, D,
the counters and the sweep count bit for bit, through the C ABI.  The
strategies here have four transitions per sweep whose first-half quads take
a pair from the previous transition's second half, so both launch plans of
a transition run."""
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
    g.set_staged(g.STAGED_AUTO)
    g.set_form(g.FORM_PLAIN)


def run(alg, G, nplus, staged, snp1="cycwor"):
    g.init_strats_builtin("cycwor", 32, snp1, G.shape[1] // 16)
    old = g.set_staged(g.STAGED_ALWAYS if staged else g.STAGED_OFF)
    try:
        r = g.hypJacL2(alg, G, nplus)
        used = g.last_staged()
    finally:
        g.set_staged(old)
    return r, used


def assert_same(a, b, ctx):
    assert a.ret == 0 and b.ret == 0, ctx
    assert a.sweeps == b.sweeps, ctx
    assert (a.rot_s, a.rot_b) == (b.rot_s, b.rot_b), ctx
    np.testing.assert_array_equal(a.D, b.D, err_msg=ctx)
    np.testing.assert_array_equal(a.G, b.G, err_msg=ctx)
    assert (a.V is None) == (b.V is None), ctx
    if a.V is not None:
        np.testing.assert_array_equal(a.V, b.V, err_msg=ctx)


@pytest.mark.parametrize("snp1", ["cycwor", "rowcyc", "colcyc", "cycloc"])
@pytest.mark.parametrize("alg,kind", [(9, "svd"), (8, "svd"), (1, "svd"), (0, "hyp"), (9, "hyp"),
                                      (9, "neg"), (11, "svd"), (11, "hyp"), (3, "hyp"), (10, "neg")])
def test_staged_equals_step_by_step(snp1, alg, kind):
    m, n = 512, 256
    if kind == "hyp":
        G, nplus = graded_input(m, n, seed=alg + 3), 100
    else:
        s = 10.0 ** (-9.0 * np.random.default_rng(alg).uniform(size=n))
        G, nplus = svd_input(m, n, s, seed=alg + 7), (n if kind == "svd" else 0)
    a, used_a = run(alg, G, nplus, True, snp1)
    b, used_b = run(alg, G, nplus, False, snp1)
    assert used_a and not used_b
    assert_same(a, b, "%s alg %d %s" % (snp1, alg, kind))


@pytest.mark.parametrize("m,n,alg", [(256, 128, 9), (1024, 512, 9), (256, 128, 11), (512, 256, 3)])
def test_staged_against_oracle(m, n, alg):
    G = graded_input(m, n, seed=n)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    s0, s1 = g.current_strats()
    old = g.set_staged(g.STAGED_ALWAYS)
    try:
        r = g.hypJacL2(alg, G, n // 3)
        assert g.last_staged()
    finally:
        g.set_staged(old)
    ro = o.hypJacL2(alg, G, n // 3, s0, s1, nthreads=8)
    assert r.sweeps == ro["sweeps"] and (r.rot_s, r.rot_b) == (ro["rot_s"], ro["rot_b"])
    np.testing.assert_array_equal(r.D, ro["D"])
    np.testing.assert_array_equal(r.G, ro["G"])
    if ro["V"] is not None:
        np.testing.assert_array_equal(r.V, ro["V"])


def test_staged_not_used_where_it_does_not_apply():
    G = graded_input(256, 128, seed=2)
    # a strategy without closed quads
    _, used = run(9, G, 128, True, snp1="mmstep")
    assert not used
    # fewer than four pairs per step
    _, used = run(9, graded_input(128, 64, seed=3), 64, True)
    assert not used
    # the default mode: small steps run step by step
    old = g.set_staged(g.STAGED_AUTO)
    try:
        g.init_strats_builtin("cycwor", 32, "cycwor", 128 // 16)
        g.hypJacL2(9, G, 128)
        assert not g.last_staged()
    finally:
        g.set_staged(old)


@pytest.mark.parametrize("form,alg,kind,snp1", [
    ("learned+split", 9, "svd", "cycwor"), ("split", 9, "svd", "cycwor"),
    # the V update of the staged launches, a mixed J (NPLUS = N/2) and a
    # strategy whose cross transitions differ from cycwor's
    ("learned+split", 11, "svd", "cycwor"), ("learned+split", 9, "hyp", "cycwor"),
    ("learned+split", 11, "hyp", "rowcyc"), ("learned+split", 9, "svd", "rowcyc"),
    # split Grams formed by the fused items as well
    ("split+fused", 9, "svd", "cycwor"), ("learned+split+fused", 11, "hyp", "cycwor")])
def test_staged_forms_equal_step_by_step(form, alg, kind, snp1):
    # skipped and split visits inside staged launches (with the learned
    # skip: the sweep's visit lists and quads reordered by the plan): the
    # same bits as the step-by-step launches of the same form
    import os
    import gpujachx_amd.learned as learned
    if "learned" in form and not os.path.exists(learned.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 1024
    s = 10.0 ** (-10.0 * np.random.default_rng(3).uniform(size=n))
    G = svd_input(n, n, s, seed=31)
    nplus = n if kind == "svd" else n // 2
    outs = []
    old_fs = g.lib().hj_set_fused_split(1 if "fused" in form else 0)
    try:
        for staged in (True, False):
            if "learned" in form:
                learned.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
                g.lib().hj_set_skip_min_visits(0)
                g.lib().hj_set_skip_min_propose(0.0)
            else:
                g.set_form(g.FORM_SPLIT_GRAM)
            try:
                r, used = run(alg, G, nplus, staged, snp1)
                outs.append((r, used, g.last_skips()[1], g.last_splits()[1]))
            finally:
                g.set_form(g.FORM_PLAIN)
                g.lib().hj_set_skip_min_visits(32000)
                g.lib().hj_set_skip_min_propose(0.1)
    finally:
        g.lib().hj_set_fused_split(old_fs)
    (a, ua, ska, spa), (b, ub, skb, spb) = outs
    assert ua and not ub
    assert (ska, spa) == (skb, spb)
    assert spa > 0
    if "learned" in form:
        assert ska > 0
    assert_same(a, b, "%s alg %d %s %s (skipped %d, split %d)" % (form, alg, kind, snp1, ska, spa))


@pytest.mark.parametrize("form,sweeps", [("plain", 1), ("learned+split", 12)])
def test_staged_full_size(form, sweeps):
    # N = 8192, the bench's size and launch plan (256 pairs per step, the
    # cross transitions where the strategy puts them): the first sweeps of
    # the staged and the step-by-step launches, every bit of G and D equal;
    # the learned form runs into the sweeps where it skips
    import ctypes
    import os
    import torch
    import gpujachx_amd.learned as learned
    if "learned" in form and not os.path.exists(learned.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 8192
    gen = torch.Generator(device="cuda").manual_seed(17)
    scale = torch.pow(10.0, -6.0 * torch.rand(n, generator=gen, device="cuda", dtype=torch.float64))
    G0 = torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64) * scale[:, None]
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    outs = []
    old_cap = g.set_max_sweeps(sweeps)
    try:
        for mode in (g.STAGED_ALWAYS, g.STAGED_OFF):
            if "learned" in form:
                learned.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
            old = g.set_staged(mode)
            try:
                W = G0.clone()
                D = torch.empty(n, dtype=torch.float64, device="cuda")
                sw = ctypes.c_uint(0)
                gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
                rc = g.lib().hypJacL2_dev(9, n, n, n, W.data_ptr(), n, None, 0, D.data_ptr(),
                                          ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), None,
                                          torch.cuda.current_stream().cuda_stream)
                torch.cuda.synchronize()
                assert rc == 0
                outs.append((W, D, sw.value, gs.value, gb.value, g.last_staged(), g.last_skips()[1]))
            finally:
                g.set_staged(old)
                g.set_form(g.FORM_PLAIN)
    finally:
        g.set_max_sweeps(old_cap)
    (Wa, Da, *ra), (Wb, Db, *rb) = outs
    assert ra[3] and not rb[3]
    assert ra[:3] == rb[:3] and ra[4] == rb[4], (ra, rb)
    if "learned" in form:
        assert ra[4] > 0, ra
    assert torch.equal(Da, Db)
    assert torch.equal(Wa, Wb)
