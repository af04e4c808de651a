"""The benchmark configurations at their full sizes.

* Hyperbolic SVD, M = 4096, N = 2048, NPLUS = 1024, graded columns, ALG 11:
  the first sweep (then the column normalisation) on the GPU equals the
  oracle's bit for bit -- U, V, D and the rotation counts -- both as
  step-by-step launches and as staged sweeps.
* SVD, M = N = 8192, the headline matrix (G = U diag(10^-12u) V^T, Haar U
  and V), run to convergence: the learned + split form skips visits, forms
  split Grams, ends on an unskipped all-FP64 sweep without a big rotation,
  and its singular values agree with the plain FP64 run's per value within
  ten times the spread of two FP64 runs that differ only in their strategy
  (cycwor against rowcyc), the large ones to 1e-11.
"""
import ctypes
import os

import numpy as np
import pytest

import gpujachx_amd as g
import gpujachx_amd.learned as lr
import oracle as o
from conftest import assert_per_value, graded_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    g.set_form(g.FORM_PLAIN)
    yield
    g.set_form(g.FORM_PLAIN)


@pytest.mark.parametrize("staged", [g.STAGED_OFF, g.STAGED_ALWAYS])
def test_hsvd_4096x2048_first_sweep_equals_oracle(staged):
    m, n, nplus = 4096, 2048, 1024
    G = graded_input(m, n, seed=2)
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    s0, s1 = g.current_strats()
    old_cap, old_stage = g.set_max_sweeps(1), g.set_staged(staged)
    try:
        print("case: hsvd %dx%d nplus %d, one sweep, staged mode %d" % (m, n, nplus, staged),
              flush=True)
        rg = g.hypJacL2(11, G, nplus)
        assert g.last_staged() == (staged == g.STAGED_ALWAYS)
    finally:
        g.set_max_sweeps(old_cap)
        g.set_staged(old_stage)
    old = o.set_max_sweeps(1)
    try:
        ro = o.hypJacL2(11, G, nplus, s0, s1, nthreads=os.cpu_count() or 8)
    finally:
        o.set_max_sweeps(old)
    assert rg.ret == 0 and ro["ret"] == 0
    assert rg.sweeps == ro["sweeps"] == 1
    assert (rg.rot_s, rg.rot_b) == (ro["rot_s"], ro["rot_b"])
    np.testing.assert_array_equal(rg.D, ro["D"])
    np.testing.assert_array_equal(rg.G, ro["G"])
    np.testing.assert_array_equal(rg.V, ro["V"])


def _svd_8192(torch, n, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)

    def haar():
        q, r = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
        return q * torch.sign(torch.diagonal(r))

    U, V = haar(), haar()
    s = torch.pow(10.0, -12.0 * torch.rand(n, dtype=torch.float64, device="cuda", generator=gen))
    return ((U * s) @ V.T).T.contiguous()  # column-major G: the row-major transpose


def _run_dev(torch, G0, alg, nplus):
    n = G0.shape[0]
    W = G0.clone()
    D = torch.empty(n, dtype=torch.float64, device="cuda")
    sw, gs, gb = ctypes.c_uint(0), ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    rc = g.lib().hypJacL2_dev(alg, n, n, nplus, W.data_ptr(), n, None, 0, D.data_ptr(),
                              ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), None,
                              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0
    return D.cpu().numpy(), sw.value


def test_svd_8192_learned_split_to_convergence():
    import torch
    if not os.path.exists(lr.DEFAULT_MODEL):
        pytest.skip("no trained model file")
    n = 8192
    G0 = _svd_8192(torch, n, seed=33)
    old_cap = g.set_max_sweeps(80)
    try:
        print("case: svd 8192 plain cycwor", flush=True)
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        Dp, swp = _run_dev(torch, G0, 9, n)
        print("case: svd 8192 plain rowcyc", flush=True)
        g.init_strats_builtin("cycwor", 32, "rowcyc", n // 16)
        Dr, _ = _run_dev(torch, G0, 9, n)
        print("case: svd 8192 learned+split", flush=True)
        g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
        lr.enable(form=g.FORM_LEARNED_SKIP | g.FORM_SPLIT_GRAM)
        Dl, swl = _run_dev(torch, G0, 9, n)
        skip_per, skipped, _ = g.last_skips()
        split_per, splits = g.last_splits()
        _, rot_b, _ = g.last_sweeps()
    finally:
        g.set_form(g.FORM_PLAIN)
        g.set_max_sweeps(old_cap)
    assert skipped > 0 and splits > 0
    # the run ends on an unskipped, all-FP64 sweep without a big rotation
    assert rot_b[-1] == 0 and skip_per[-1] == 0 and split_per[-1] == 0
    assert swl <= swp + 2, (swl, swp)
    sp, sr, sl = (np.sqrt(np.abs(x)) for x in (Dp, Dr, Dl))
    assert_per_value(sl, sp, sr, ctx="svd 8192 learned+split")
