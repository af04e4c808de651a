"""The thin level-1 layer (share/device_code.hpp:6-21: initSymbols,
defJacL1{,v,s,sv}, hypJacL1{,v,s,sv}, initD, initV) driven step by step the
way the reference's own level-2 loop drives it, against the oracle run for
the same number of steps: U, V, D and the convergence counter bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


def _bind():
    L = g.lib()
    u, i, vp = ctypes.c_uint, ctypes.c_int, ctypes.c_void_p
    L.initSymbols.argtypes = [vp, vp, vp, u, u, u, u, u]
    L.initSymbols.restype = None
    for name in ("defJacL1", "defJacL1v", "defJacL1s", "defJacL1sv"):
        getattr(L, name).argtypes = [u, i]
        getattr(L, name).restype = None
    for name in ("hypJacL1", "hypJacL1v", "hypJacL1s", "hypJacL1sv"):
        getattr(L, name).argtypes = [u, u]
        getattr(L, name).restype = None
    L.initD.argtypes = [vp, vp, u, u, u, u, u]
    L.initD.restype = None
    L.initV.argtypes = [vp, u, u, u]
    L.initV.restype = None
    return L


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    g.set_factorization(g.FACT_CHOLESKY)


# (ALG, definite / hyperbolic): the kernel family is the one the reference's
# 1/HypJacL2.cu:110-113 picks for that ALG
@pytest.mark.parametrize("alg,kind,fact", [
    (8, "svd", 0), (9, "svd", 0), (11, "svd", 0), (10, "svd", 0),
    (11, "hyp", 0), (9, "hyp", 0), (1, "svd", 0), (3, "hyp", 0),
    (9, "svd", 1), (11, "hyp", 2),
])
def test_thin_layer_one_sweep(alg, kind, fact):
    L = _bind()
    m, n = 256, 128
    if kind == "svd":
        G = svd_input(m, n, 10.0 ** (-8.0 * np.random.default_rng(alg).uniform(size=n)), seed=alg)
        nplus = n
    else:
        G = graded_input(m, n, seed=alg)
        nplus = 50
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    s0, s1 = g.current_strats()
    steps = s1.shape[0]
    sort, full, blk = bool(alg & 1), bool(alg & 2), bool(alg & 8)
    nswp = 1 if blk else 30
    Gd = torch.from_numpy(np.asfortranarray(G).T.copy()).cuda()          # column-major, ld = m
    Vd = torch.zeros((n, n), dtype=torch.float64, device="cuda") if full else None
    Dd = torch.zeros(n, dtype=torch.float64, device="cuda")
    cvg = torch.zeros(1, dtype=torch.int64, device="cuda")
    g.set_factorization(fact)
    try:
        L.initSymbols(Gd.data_ptr(), Vd.data_ptr() if full else None, cvg.data_ptr(), m, n, m,
                      n if full else 0, nswp)
        if full:
            L.initV(Vd.data_ptr(), 0, n, n)
        name = ("def" if nplus == n else "hyp") + "JacL1" + ("s" if sort else "") + ("v" if full else "")
        fn = getattr(L, name)
        for st in range(steps):
            fn(st, 1 if nplus == n else nplus)
        L.initD(Gd.data_ptr(), Dd.data_ptr(), 0, m, n, nplus, m)
        torch.cuda.synchronize()
    finally:
        g.set_factorization(g.FACT_CHOLESKY)
    old = o.set_factorization(fact)
    try:
        ro = o.hypJacL2(alg, G, nplus, s0, s1, max_steps=steps)
    finally:
        o.set_factorization(old)
    ctx = "%s alg %d fact %d" % (name, alg, fact)
    np.testing.assert_array_equal(Gd.cpu().numpy().T, ro["G"], err_msg=ctx)
    np.testing.assert_array_equal(Dd.cpu().numpy(), ro["D"], err_msg=ctx)
    if full:
        np.testing.assert_array_equal(Vd.cpu().numpy().T, ro["V"], err_msg=ctx)
    c = int(cvg.cpu().numpy()[0]) & 0xFFFFFFFFFFFFFFFF
    assert (c & 0xFFFFFFFF, c >> 32) == (ro["rot_s_trace"][0], ro["rot_b_trace"][0]), ctx
