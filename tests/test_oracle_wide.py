"""The oracle of the wide-block form (blocks of 32 columns, 64-column pairs),
pinned on the reference's width (w = 16: the same bits as the pinned w = 16
oracle for every ALG and J kind), by exact integer Grams, by the invariants
of a visit (W orthogonal / J-orthogonal, the pair orthogonalised), the
update against a matrix product, and whole runs against LAPACK and the
closed-form spectrum."""
import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input


def _tables(width, n):
    s0 = g.strat_generate("cycwor", 2 * width).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // width)
    return s0, s1


@pytest.mark.parametrize("alg", [0, 1, 2, 3, 8, 9, 10, 11])
@pytest.mark.parametrize("nplus", [128, 0, 40, 64, 100])
def test_width16_equals_reference_width_oracle(alg, nplus):
    # the generic code at the reference's width: every bit of G, V, D and the
    # counters as the w = 16 oracle
    m, n = 256, 128
    G = graded_input(m, n, seed=3 + nplus, lo=-3, hi=3)
    s0, s1 = _tables(16, n)
    a = o.hypJacL2(alg, G, nplus, s0, s1, nthreads=4)
    b = o.hypJacL2_w(alg, G, nplus, 16, s0, s1, max_sweeps=30)
    assert a["ret"] == 0 and b["ret"] == 0
    assert (a["sweeps"], a["rot_s"], a["rot_b"]) == (b["sweeps"], b["rot_s"], b["rot_b"])
    assert a["rot_b_trace"] == b["rot_b_trace"]
    for k in ("G", "D") + (("V",) if alg & 2 else ()):
        np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64), err_msg=k)


@pytest.mark.parametrize("nrow", [64, 256, 1024])
def test_wide_gram_exact_on_small_integers(nrow):
    # integers |x| <= 15 and power-of-two column scales: every product and
    # partial sum is exact, so the tiles must equal G^T G entry for entry
    # (pins the tile order, lane layout and row coverage)
    rng = np.random.default_rng(nrow)
    G = rng.integers(-15, 16, (nrow, 128)).astype(np.float64)
    G[:, 37] *= 2.0 ** 30
    G[:, 70] *= 2.0 ** -30
    for a, b in ((0, 1), (3, 1), (2, 0)):
        p = o.gram_partial_w(G, 32, a, b, 0, nrow)
        A = o.tiles_to_lower(p, 32)
        X = np.concatenate([G[:, 32 * a:32 * a + 32], G[:, 32 * b:32 * b + 32]], axis=1)
        # tiles ti >= tj (the diagonal tiles whole)
        full = X.T @ X
        mask = np.kron(np.tril(np.ones((4, 4))), np.ones((16, 16))) > 0
        np.testing.assert_array_equal(A[mask], full[mask])
        assert not A[~mask].any()


def test_wide_gram_chunks_add_up():
    # partials of row chunks add to the Gram of all rows (relative rounding)
    rng = np.random.default_rng(2)
    G = rng.standard_normal((512, 64)) * 10.0 ** rng.uniform(-3, 3, 64)
    whole = o.tiles_to_lower(o.gram_partial_w(G, 32, 0, 1, 0, 512), 32)
    parts = sum(o.tiles_to_lower(o.gram_partial_w(G, 32, 0, 1, r0, 128), 32)
                for r0 in range(0, 512, 128))
    full = G.T @ G
    mask = np.kron(np.tril(np.ones((4, 4))), np.ones((16, 16))) > 0
    d = np.sqrt(np.diag(full))
    for A in (whole, parts):
        assert (np.abs(A - full) / np.outer(d, d))[mask].max() < 1e-14


@pytest.mark.parametrize("sort", [0, 1])
def test_wide_visit_orthogonalises_the_pair(sort):
    # a visit with the inner sweeps run to convergence: W orthogonal, the
    # pair's columns G W mutually orthogonal; with sort, norms descending
    rng = np.random.default_rng(5 + sort)
    X = rng.standard_normal((256, 64)) * 10.0 ** rng.uniform(-2, 2, 64)
    s0, _ = _tables(32, 64)
    p = o.gram_partial_w(X, 32, 0, 1, 0, 256)
    W, s, b, upd = o.pair_w(p, 32, sort, 0, 1, 0, s0, nswp=30)
    assert s > 0 and b > 0 and upd == 1
    assert np.abs(W.T @ W - np.eye(64)).max() < 1e-13
    Y = X @ W
    nrm = np.linalg.norm(Y, axis=0)
    C = (Y.T @ Y) / np.outer(nrm, nrm)
    assert np.abs(C - np.eye(64)).max() < 1e-12
    if sort:
        assert np.all(np.diff(nrm) <= 0)


@pytest.mark.parametrize("npos", [0, 20, 32, 50, 64])
def test_wide_hyperbolic_visit(npos):
    # a pair mixing J signs (local npos): W is J-orthogonal and the pair's
    # columns G W J-orthogonal; the first npos columns of J positive
    rng = np.random.default_rng(npos)
    X = rng.standard_normal((256, 64)) * 10.0 ** rng.uniform(-1, 1, 64)
    s0, _ = _tables(32, 64)
    p = o.gram_partial_w(X, 32, 0, 1, 0, 256)
    definite = 1 if npos == 64 else (-1 if npos == 0 else 0)
    hyp = 1 if definite == 0 else 0
    J = np.diag(np.r_[np.ones(npos), -np.ones(64 - npos)])
    for sort in (0, 1):
        W, s, b, upd = o.pair_w(p, 32, sort, hyp, definite, npos, s0, nswp=30)
        assert s > 0 and upd == 1
        assert np.abs(W.T @ J @ W - J).max() < 1e-11 * max(1.0, np.abs(W).max() ** 2)
        Y = X @ W
        nrm = np.linalg.norm(Y, axis=0)
        C = (Y.T @ Y) / np.outer(nrm, nrm)
        assert np.abs(C - np.eye(64)).max() < 1e-11


def test_wide_update_is_the_product():
    rng = np.random.default_rng(8)
    G = rng.standard_normal((128, 128))
    Wm = rng.standard_normal((64, 64))
    out = o.update_w(G, 32, 3, 1, Wm)
    X = np.concatenate([G[:, 96:128], G[:, 32:64]], axis=1)
    Y = X @ Wm
    bound = 64 * 2.0 ** -52 * (np.abs(X) @ np.abs(Wm))
    assert np.all(np.abs(out[:, 96:128] - Y[:, :32]) <= bound[:, :32])
    assert np.all(np.abs(out[:, 32:64] - Y[:, 32:]) <= bound[:, 32:])
    np.testing.assert_array_equal(out[:, :32], G[:, :32])
    np.testing.assert_array_equal(out[:, 64:96], G[:, 64:96])
    # a permutation moves columns exactly, across the two blocks
    P = np.eye(64)[:, rng.permutation(64)]
    outp = o.update_w(G, 32, 3, 1, P)
    np.testing.assert_array_equal(np.concatenate([outp[:, 96:128], outp[:, 32:64]], axis=1), X @ P)


@pytest.mark.parametrize("alg", [9, 11, 1, 3])
def test_wide_svd_closed_form_spectrum(alg):
    n = 256
    s = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    G = svd_input(n, n, s, seed=1)
    s0, s1 = _tables(32, n)
    r = o.hypJacL2_w(alg, G, n, 32, s0, s1)
    assert r["ret"] == 0 and r["rot_b_trace"][-1] == 0
    sig = np.sqrt(r["D"])
    np.testing.assert_allclose(np.sort(sig)[::-1], s, rtol=1e-9)
    U = r["G"]
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    if alg & 2:
        V = r["V"]
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
        assert np.abs(G @ V - U * sig).max() < 1e-13
    if alg & 1:
        for b in range(n // 32):
            assert np.all(np.diff(r["D"][32 * b:32 * b + 32]) <= 0)


def test_wide_vs_dgesvj():
    from oracle.lapack import dgesvj
    rng = np.random.default_rng(9)
    m, n = 512, 256
    G = rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-6, 6, n)
    s0, s1 = _tables(32, n)
    r = o.hypJacL2_w(11, G, n, 32, s0, s1)
    sj, _, _, info, _ = dgesvj(G, want_v=False)
    assert info == 0
    np.testing.assert_allclose(np.sort(np.sqrt(r["D"]))[::-1], np.sort(sj)[::-1], rtol=1e-12)


@pytest.mark.parametrize("npl", [48, 96, 130])
def test_wide_hsvd_invariants(npl):
    m, n = 256, 256
    G = graded_input(m, n, seed=12, lo=-3, hi=3)
    s0, s1 = _tables(32, n)
    r = o.hypJacL2_w(11, G, npl, 32, s0, s1)
    assert r["ret"] == 0 and r["rot_b_trace"][-1] == 0
    J = np.diag(np.r_[np.ones(npl), -np.ones(n - npl)])
    V, U, D = r["V"], r["G"], r["D"]
    assert np.abs(V.T @ J @ V - J).max() < 1e-10 * np.abs(V).max() ** 2
    assert np.abs(G @ V - U * np.sqrt(np.abs(D))).max() < 1e-12 * np.abs(G).max() * np.abs(V).max()
    assert np.all(D[:npl] > 0) and np.all(D[npl:] < 0)
    Vinv = J @ V.T @ J
    Gr = (U * np.sqrt(np.abs(D))) @ Vinv
    assert np.abs(Gr - G).max() < 1e-11 * np.abs(G).max()
    # the same spectrum as the method at the reference's width
    s0n, s1n = _tables(16, n)
    q = o.hypJacL2(11, G, npl, s0n, s1n, nthreads=4)
    np.testing.assert_allclose(np.sort(D), np.sort(q["D"]), rtol=1e-11)


@pytest.mark.parametrize("args,code", [((9, 256, 96, 96), -3), ((9, 256, 32, 32), -3),
                                       ((16, 256, 64, 64), -1), ((9, 256, 64, 65), -4)])
def test_wide_argument_codes(args, code):
    alg, m, n, npl = args
    s0, s1 = _tables(32, 64)
    G = np.zeros((m, n))
    assert o.hypJacL2_w(alg, G, npl, 32, s0, s1)["ret"] == code
