"""Dataflow sweeps (one persistent launch per sweep, pair visits overlapped
with other pairs' block updates) against the step-by-step launches: the
same U, D, sweep count and rotation counts, bit for bit, for every routine
and strategy the dataflow path takes, and the step-by-step path where it
does not apply."""
import numpy as np
import pytest

import gpujachx_amd as g
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_dataflow(False)
    g.set_form(g.FORM_PLAIN)


def _pair(alg, G, nplus, strat="cycwor", form=None):
    n = G.shape[1]
    g.init_strats_builtin("cycwor", 32, strat, n // 16)
    out = []
    for flow in (False, True):
        g.set_dataflow(flow)
        if form is not None:
            form()
        try:
            r = g.hypJacL2(alg, G, nplus)
            used = g.last_dataflow()
            sw = g.last_sweeps()
            sk = g.last_skips()
        finally:
            g.set_dataflow(False)
            g.set_form(g.FORM_PLAIN)
        assert r.ret == 0
        out.append((r, used, sw, sk))
    return out


def _same(a, b):
    ra, rb = a[0], b[0]
    assert ra.sweeps == rb.sweeps
    assert (ra.rot_s, ra.rot_b) == (rb.rot_s, rb.rot_b)
    assert a[2][:2] == b[2][:2]
    np.testing.assert_array_equal(ra.G, rb.G)
    np.testing.assert_array_equal(ra.D, rb.D)


@pytest.mark.parametrize("alg,m,n,kind,strat", [
    (9, 512, 256, "svd", "cycwor"),
    (8, 512, 256, "svd", "cycwor"),
    (9, 256, 256, "hyp", "cycwor"),
    (1, 256, 128, "svd", "cycwor"),
    (0, 256, 128, "hyp", "cycwor"),
    (9, 512, 256, "svd", "rowcyc"),
    (9, 512, 256, "svd", "colcyc"),
    (9, 512, 256, "svd", "cycloc"),
    (9, 64, 64, "svd", "cycwor"),
    (9, 2048, 1024, "svd", "cycwor"),
    (9, 4096, 2048, "hyp", "cycwor"),
])
def test_flow_matches_steps(alg, m, n, kind, strat):
    if kind == "svd":
        G = svd_input(m, n, 10.0 ** (-10.0 * np.random.default_rng(n).uniform(size=n)), seed=m + n)
        nplus = n
    else:
        G = graded_input(m, n, seed=m + n)
        nplus = n // 3
    steps, flow = _pair(alg, G, nplus, strat)
    assert not steps[1] and flow[1], "dataflow path not taken"
    _same(steps, flow)


@pytest.mark.parametrize("strat", ["BrentL", "mmstep"])
def test_flow_not_taken_without_closed_quads(strat):
    G = graded_input(256, 128, seed=3)
    steps, flow = _pair(9, G, 128, strat)
    assert not flow[1]
    _same(steps, flow)


def test_flow_not_taken_with_v():
    G = graded_input(256, 128, seed=4)
    steps, flow = _pair(11, G, 64)
    assert not flow[1]
    _same(steps, flow)
    np.testing.assert_array_equal(steps[0].V, flow[0].V)


def test_flow_split_form_matches_steps():
    n = 512
    G = svd_input(1024, n, 10.0 ** (-10.0 * np.random.default_rng(9).uniform(size=n)), seed=9)
    steps, flow = _pair(9, G, n, form=lambda: g.set_form(g.FORM_SPLIT_GRAM))
    assert flow[1]
    _same(steps, flow)


def test_flow_learned_split_matches_steps():
    import os
    import gpujachx_amd.learned as lr
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 1024
    G = svd_input(1024, n, 10.0 ** (-12.0 * np.random.default_rng(2).uniform(size=n)), seed=43)
    old = g.set_max_sweeps(80)

    def form():
        lr.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
        g.lib().hj_set_skip_min_visits(0)
    try:
        steps, flow = _pair(9, G, n, form=form)
    finally:
        g.set_max_sweeps(old)
        g.lib().hj_set_skip_min_visits(32000)
    assert flow[1]
    assert steps[3][1] > 0, "no visit was skipped"
    assert steps[3][:2] == flow[3][:2]
    _same(steps, flow)


def test_flow_profile_counts_one_launch_per_sweep():
    G = graded_input(512, 256, seed=6)
    g.init_strats_builtin("cycwor", 32, "cycwor", 16)
    g.set_dataflow(True)
    g.set_profile(True)
    try:
        r = g.hypJacL2(9, G, 256)
        ms, cnt = g.kernel_ms(3)
        used = g.last_dataflow()
    finally:
        g.set_profile(False)
        g.set_dataflow(False)
    assert used and r.ret == 0
    assert cnt == r.sweeps and ms > 0.0
