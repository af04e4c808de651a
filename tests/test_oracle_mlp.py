"""Pins the CPU MLP restatement (oracle/mlp.py) by mathematics alone:
central finite differences of the loss, a hand-computed tiny case, bf16
rounding against its definition."""
import numpy as np

from oracle import mlp


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(mlp.NPARAM) * scale


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((40, mlp.F))
    y = (rng.uniform(size=40) < 0.3).astype(np.float64)
    p = _params(2)
    _, g = mlp.loss_and_grad(X, y, p, wpos=3.0)
    h = 1e-6
    idx = np.r_[rng.choice(mlp.NPARAM, 300, replace=False), mlp.OFF_B3, mlp.OFF_B1, mlp.OFF_B2]
    for i in idx:
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        lp, _ = mlp.loss_and_grad(X, y, p + e, 3.0)
        lm, _ = mlp.loss_and_grad(X, y, p - e, 3.0)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_forward_hand_case():
    # one input feature active, two hidden units active, everything else 0
    p = np.zeros(mlp.NPARAM)
    p[mlp.OFF_W1 + 0 * mlp.H + 0] = 2.0      # h1_0 = relu(2 x0)
    p[mlp.OFF_W1 + 0 * mlp.H + 1] = -1.0     # h1_1 = relu(-x0)
    p[mlp.OFF_W2 + 0 * mlp.H + 5] = 1.5      # h2_5 = relu(1.5 h1_0 + 0.25)
    p[mlp.OFF_B2 + 5] = 0.25
    p[mlp.OFF_W3 + 5] = -2.0
    p[mlp.OFF_B3] = 0.5
    X = np.zeros((2, mlp.F))
    X[0, 0], X[1, 0] = 1.0, -1.0
    z, _ = mlp.forward(X, p)
    assert z[0] == 0.5 - 2.0 * (1.5 * 2.0 + 0.25)
    assert z[1] == 0.5 - 2.0 * 0.25
    # weighted BCE at z = 0 is wt * log 2
    loss, _ = mlp.loss_and_grad(np.zeros((3, mlp.F)), np.array([1.0, 0.0, 1.0]), np.zeros(mlp.NPARAM), 4.0)
    assert abs(loss - (4 + 1 + 4) / 3 * np.log(2.0)) < 1e-15


def test_bf16_round_definition():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -3.14159, 1e-40, np.inf],
                 dtype=np.float32)
    r = mlp.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                            # tie at 2^-8 rounds to even (1.0)
    assert r[2] == 1.0                            # below the tie
    assert r[3] == np.float32(1.0 + 2 ** -7)      # above the tie
    assert abs(r[4] - -3.140625) == 0.0
    assert np.isinf(r[6])
    # every rounded value has 16 zero low bits and is the nearest such value
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(10000) * 10.0 ** rng.uniform(-20, 20, 10000)).astype(np.float32)
    rx = mlp.bf16_round(x)
    assert np.all((rx.view(np.uint32) & 0xFFFF) == 0)
    up = (rx.view(np.uint32) + 0x10000).view(np.float32)
    dn = (rx.view(np.uint32) - 0x10000).view(np.float32)
    err = np.abs(rx.astype(np.float64) - x)
    assert np.all(err <= np.abs(up.astype(np.float64) - x))
    assert np.all(err <= np.abs(dn.astype(np.float64) - x))


def test_features_definitions():
    rec = np.zeros(3, dtype=mlp.REC_DTYPE)
    rec["rot_b"] = [0, 10, 1000]
    rec["rot_s"] = [5, 20, 1000]
    rec["maxcos"] = [1e-8, 0.5, 0.0]
    rec["l2ni"] = [3.0, -2.0, 0.0]
    rec["l2nj"] = [1.0, 2.0, 0.0]
    rec["sweep"] = [4, 3, 0]
    rec["skips"] = [0, 1, 0]
    s1 = np.array([[0, 1], [2, 3], [4, 5]], dtype=np.uint16)
    X = mlp.features(rec, s1, nplus=40, n1=8, sweep_now=5, prev_frac=1e-3)
    assert X[0, 0] == 0 and X[1, 0] == 10 / 496 and X[2, 0] == 1.0
    assert abs(X[0, 2] - (-8 / 16)) < 1e-7
    assert X[0, 3] == 2 / 64 and X[1, 4] == 4 / 64
    assert X[0, 6] == 1 / 4 and X[1, 6] == 2 / 4 and X[1, 7] == 1 / 4
    assert abs(X[0, 9] - (-3 / 12)) < 1e-12
    # nplus = 40: blocks 0,1 positive; block 2 has 8 negative columns; 3.. negative
    assert list(X[:, 10]) == [0.0, (8 + 16) / 32, 1.0]
    assert X[0, 11] == 3 / 10
