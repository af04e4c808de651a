"""CPU restatement of the pair-skip MLP (features, forward, weighted BCE
loss, backward) in float64 -- TEST INFRASTRUCTURE ONLY, the checker of the
library's HIP training and inference kernels (gpujachx_amd/csrc/hj_mlp.hip).

The learned skip has no counterpart in the reference; what fixes these
formulas is the mathematics: the backward pass is pinned by central finite
differences of the loss (tests/test_oracle_mlp.py), the forward pass by a
hand-computed two-unit case, the features by their definitions.

Parameter layout: W1[F][H], b1[H], W2[H][H], b2[H], W3[H], b3 (F=12, H=64).
The layers use the bf16-rounded parameters; the gradient is that of the loss
at the rounded parameters (straight-through for the master copy).
"""
import numpy as np

F, H = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = np.where(finite, r, u.astype(np.uint32)).astype(np.uint32)
    return out.view(np.float32).reshape(np.shape(x))


def unpack(p, hidden=H, nfeat=F):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(nfeat, hidden)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(hidden, hidden)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, W3, b3


def forward(X, p):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1 + b1
    A1 = np.maximum(Z1, 0.0)
    Z2 = A1 @ W2 + b2
    A2 = np.maximum(Z2, 0.0)
    z = A2 @ W3 + b3
    return z, (X, Z1, A1, Z2, A2)


def loss_and_grad(X, y, p, wpos):
    """Mean weighted BCE with logits and its gradient w.r.t. p (float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z, (X, Z1, A1, Z2, A2) = forward(X, p)
    y = np.asarray(y, dtype=np.float64)
    B = X.shape[0]
    wt = np.where(y > 0.5, wpos, 1.0)
    sp = np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))
    loss = np.mean(wt * (sp - y * z))
    dz = wt * (1.0 / (1.0 + np.exp(-z)) - y) / B
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = A2.T @ dz
    g[OFF_B3] = dz.sum()
    dZ2 = np.outer(dz, W3) * (Z2 > 0)
    g[OFF_W2:OFF_B2] = (A1.T @ dZ2).reshape(-1)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    g[OFF_W1:OFF_B1] = (X.T @ dZ1).reshape(-1)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    return loss, g


def predict(X, p):
    z, _ = forward(X, p)
    return 1.0 / (1.0 + np.exp(-z))


REC_DTYPE = np.dtype([("rot_s", "<u4"), ("rot_b", "<u4"), ("maxcos", "<f4"), ("l2ni", "<f4"),
                      ("l2nj", "<f4"), ("sweep", "<u4"), ("skips", "<u4"), ("pad", "<u4")])


def jmix(bi, bj, nplus):
    """Fraction of J = -1 columns among the 32 columns of blocks bi, bj."""
    def neg(b):
        lo = 16 * np.asarray(b, dtype=np.int64)
        return np.clip(lo + 16 - nplus, 0, 16)
    return (neg(bi) + neg(bj)) / 32.0


def features(rec, strat1, nplus, n1, sweep_now, prev_frac):
    """Feature rows of the next visit of every record (definitions of
    hj_mlp.hip visit_features), float64."""
    r = np.asarray(rec).reshape(-1)
    s = np.asarray(strat1).reshape(-1, 2).astype(np.int64)
    X = np.zeros((r.size, F))
    X[:, 0] = np.minimum(r["rot_b"], 496) / 496.0
    X[:, 1] = np.minimum(r["rot_s"], 496) / 496.0
    X[:, 2] = np.log10(np.maximum(r["maxcos"].astype(np.float64), 1e-30)) / 16.0
    d = r["l2ni"].astype(np.float64) - r["l2nj"]
    X[:, 3] = d / 64.0
    X[:, 4] = np.abs(d) / 64.0
    X[:, 5] = sweep_now / 32.0
    X[:, 6] = (sweep_now - r["sweep"].astype(np.float64)) / 4.0
    X[:, 7] = r["skips"] / 4.0
    X[:, 8] = prev_frac
    X[:, 9] = np.log10(max(prev_frac, 1e-12)) / 12.0
    X[:, 10] = jmix(s[:, 0], s[:, 1], nplus)
    X[:, 11] = np.log2(n1) / 10.0
    return X
