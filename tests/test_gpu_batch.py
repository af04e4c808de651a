"""Many uneven (H)SVDs in one persistent launch: every problem bit for bit
against the oracle (single-wave Gram order), and the launch shapes /
strategies the autotuner may pick."""
import numpy as np
import pytest
import torch

import gpujachx_amd as g
import oracle as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    g.configureGPU(0)
    yield
    o.set_single_wave_gram(False)


def _problems(seed, shapes):
    rng = np.random.default_rng(seed)
    out = []
    for (m, n, nplus, alg) in shapes:
        G = rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-3, 3, n)
        out.append((G, nplus, alg))
    return out


def _run_batch(probs, cfg):
    dev = []
    descs = []
    for G, nplus, alg in probs:
        m, n = G.shape
        Gd = torch.from_numpy(np.asfortranarray(G).T.copy()).cuda()   # column-major, ld = m
        Dd = torch.zeros(n, dtype=torch.float64, device="cuda")
        Vd = torch.zeros((n, n), dtype=torch.float64, device="cuda") if alg & 2 else None
        dev.append((Gd, Vd, Dd))
        descs.append(g.BatchDesc(alg, m, n, nplus, m, n if Vd is not None else 0, Gd.data_ptr(),
                                 Vd.data_ptr() if Vd is not None else None, Dd.data_ptr()))
    b = g.Batch(descs, cfg)
    b.run()
    torch.cuda.synchronize()
    res = b.results()
    b.close()
    outs = []
    for (Gd, Vd, Dd) in dev:
        outs.append((Gd.cpu().numpy().T, None if Vd is None else Vd.cpu().numpy().T, Dd.cpu().numpy()))
    return outs, res


def _oracle(G, nplus, alg, fam, max_sweeps):
    n = G.shape[1]
    s1 = g.strat_generate(fam, n // 16)
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    o.set_single_wave_gram(True)
    old = o.set_max_sweeps(max_sweeps)
    try:
        return o.hypJacL2(alg, G, nplus, s0, s1)
    finally:
        o.set_max_sweeps(old)
        o.set_single_wave_gram(False)


SHAPES = [(64, 64, 64, 9), (128, 64, 32, 11), (192, 128, 128, 11), (256, 192, 96, 9),
          (128, 128, 0, 1), (320, 256, 200, 3), (64, 32, 32, 11), (512, 320, 160, 11),
          (256, 256, 256, 8), (384, 192, 0, 10), (192, 96, 48, 2), (128, 96, 96, 0)]


@pytest.mark.parametrize("waves,wg,strategy", [(4, 2, 0), (1, 4, 1), (8, 1, 2), (2, 1, 0)])
def test_batch_each_problem_bitwise(waves, wg, strategy):
    probs = _problems(10 + waves, SHAPES)
    cfg = g.default_batch_cfg()
    cfg.waves, cfg.wg_per_cu, cfg.strategy, cfg.max_sweeps = waves, wg, strategy, 30
    outs, res = _run_batch(probs, cfg)
    for k, ((G, nplus, alg), (Gg, Vg, Dg)) in enumerate(zip(probs, outs)):
        n1 = G.shape[1] // 16
        fam = {0: "cycwor", 1: "BrentL", 2: "mmstep"}[strategy]
        if fam == "cycwor" and n1 & (n1 - 1):
            fam = "BrentL"
        ro = _oracle(G, nplus, alg, fam, 30)
        ctx = "problem %d %s alg %d fam %s" % (k, G.shape, alg, fam)
        assert res["sweeps"][k] == ro["sweeps"], ctx
        assert res["rot_s"][k] == ro["rot_s"] and res["rot_b"][k] == ro["rot_b"], ctx
        np.testing.assert_array_equal(Dg, ro["D"], err_msg=ctx)
        np.testing.assert_array_equal(Gg, ro["G"], err_msg=ctx)
        if alg & 2:
            np.testing.assert_array_equal(Vg, ro["V"], err_msg=ctx)
    assert np.all(res["end_us"] >= res["start_us"])


def test_batch_many_small_problems_converge():
    rng = np.random.default_rng(3)
    shapes = []
    for _ in range(300):
        n = 32 * int(rng.integers(1, 9))
        m = n * int(rng.integers(1, 3))
        m = ((m + 63) // 64) * 64
        nplus = n if rng.uniform() < 0.5 else n // 2
        shapes.append((m, n, nplus, 11))
    probs = _problems(4, shapes)
    cfg = g.default_batch_cfg()
    cfg.max_sweeps = 40
    outs, res = _run_batch(probs, cfg)
    for (G, nplus, alg), (Gg, Vg, Dg) in zip(probs, outs):
        n = G.shape[1]
        J = np.r_[np.ones(nplus), -np.ones(n - nplus)]
        # V^T J V = J and G V = U diag(sqrt|D|)
        assert np.abs(Vg.T @ (J[:, None] * Vg) - np.diag(J)).max() < 1e-9 * max(1.0, np.abs(Vg).max() ** 2)
        res_ = G @ Vg - Gg * np.sqrt(np.abs(Dg))
        assert np.abs(res_).max() < 1e-11 * np.abs(G).max() * n
    assert np.all(res["rot_b"] >= 0) and res["sweeps"].max() < 40


def test_batch_autotune_returns_valid_cfg():
    probs = _problems(6, SHAPES * 4)
    dev, descs = [], []
    for G, nplus, alg in probs:
        m, n = G.shape
        Gd = torch.from_numpy(np.asfortranarray(G).T.copy()).cuda()
        Dd = torch.zeros(n, dtype=torch.float64, device="cuda")
        Vd = torch.zeros((n, n), dtype=torch.float64, device="cuda") if alg & 2 else None
        dev.append((Gd.clone(), Gd, Vd, Dd))
        descs.append(g.BatchDesc(alg, m, n, nplus, m, n if Vd is not None else 0, Gd.data_ptr(),
                                 Vd.data_ptr() if Vd is not None else None, Dd.data_ptr()))
    best, rate = g.batch_autotune(descs, sample=24)
    assert rate > 0 and 1 <= best.waves <= 8 and best.wg_per_cu >= 1
    for G0, Gd, _, _ in dev:             # inputs untouched
        assert torch.equal(G0, Gd)
