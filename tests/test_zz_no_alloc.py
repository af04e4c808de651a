"""The allocation-free entry point (hypJacL2_dev_noalloc) against
hypJacL2_dev through the C ABI, bit for bit: two calls in a row, the second
on a smaller problem (the scratch array keeps the first call's contents),
and the rejection of a problem whose work buffers exceed the array."""
import ctypes

import numpy as np
import pytest

import gpujachx_amd as g
from conftest import graded_input, svd_input

pytestmark = pytest.mark.gpu


def run(entry, alg, G, nplus):
    import torch
    m, n = G.shape
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    dG = torch.from_numpy(np.asfortranarray(G).T.copy()).cuda()  # column-major: row-major G^T
    dV = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    dD = torch.zeros(n, dtype=torch.float64, device="cuda")
    sw, gs, gb = ctypes.c_uint(0), ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    stream = torch.cuda.current_stream().cuda_stream
    print("case: %s alg %d m %d n %d nplus %d" % (entry, alg, m, n, nplus), flush=True)
    rc = getattr(g.lib(), entry)(alg, m, n, nplus, dG.data_ptr(), m, dV.data_ptr(), n,
                                 dD.data_ptr(), ctypes.byref(sw), ctypes.byref(gs),
                                 ctypes.byref(gb), None, stream)
    torch.cuda.synchronize()
    return rc, dG.cpu().numpy(), dV.cpu().numpy(), dD.cpu().numpy(), sw.value, gs.value, gb.value


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    g.set_form(g.FORM_PLAIN)
    g.set_step_limit(0)
    yield


def test_noalloc_equals_hypjacl2_dev_twice():
    cases = [
        (11, svd_input(1024, 512, 10.0 ** (-8.0 * np.arange(512) / 511), seed=71), 512),
        (9, graded_input(256, 128, seed=72), 48),  # the second call is smaller
    ]
    for alg, G, nplus in cases:
        ref = run("hypJacL2_dev", alg, G, nplus)
        out = run("hypJacL2_dev_noalloc", alg, G, nplus)
        assert ref[0] == 0 and out[0] == 0
        assert out[4:] == ref[4:], (alg, out[4:], ref[4:])
        np.testing.assert_array_equal(out[1], ref[1])
        np.testing.assert_array_equal(out[3], ref[3])
        if alg & 2:
            np.testing.assert_array_equal(out[2], ref[2])


def test_noalloc_rejects_work_beyond_the_array():
    import torch
    n = 4096  # 128 pairs x 16 Gram chunks of partials: more than the 8 MiB array
    g.init_strats_builtin("cycwor", 32, "cycwor", n // 16)
    dG = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    dD = torch.zeros(n, dtype=torch.float64, device="cuda")
    sw, gs, gb = ctypes.c_uint(0), ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    rc = g.lib().hypJacL2_dev_noalloc(9, n, n, n, dG.data_ptr(), n, None, 0, dD.data_ptr(),
                                      ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), None,
                                      torch.cuda.current_stream().cuda_stream)
    assert rc == -14  # HJ_ERR_SCRATCH
