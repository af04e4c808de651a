"""The oracle's QR path and Drmac level-0 (the reference's USE_QR /
USE_DRMAC options) pinned by what the mathematics fixes: norms against exact
rational arithmetic at every scale, the Givens rotation's defining
equations, the R factor against RᵀR = GᵀG and numpy's |R|, and whole runs
against the constructed spectrum and the Cholesky path."""
import math
from fractions import Fraction

import numpy as np
import pytest

import gpujachx_amd as g
import oracle as o
from conftest import graded_input, svd_input


def exact_norm(x):
    """sqrt(sum x_i^2) correctly rounded, from exact rationals."""
    s = sum(Fraction(float(v)) ** 2 for v in x)
    if s == 0:
        return 0.0
    # integer square root of s scaled by 4^k, k chosen for > 60 bits
    k = max(0, 120 - (s.numerator.bit_length() - s.denominator.bit_length()) // 2)
    num = s.numerator * (1 << (2 * k))
    r = math.isqrt(num // s.denominator)
    return math.ldexp(float(r), -k)


@pytest.mark.parametrize("scale", [1.0, 1e-300, 1e-160, 1e-100, 1e100, 1e160, 1e300])
def test_nrm2_32_every_scale(scale):
    rng = np.random.default_rng(int(abs(math.log10(scale))) + 1)
    for _ in range(20):
        x = rng.standard_normal(32) * scale
        ref = exact_norm(x)
        got = o.nrm2_32(x)
        assert np.isfinite(got)
        assert abs(got - ref) <= 4e-16 * ref, (scale, got, ref)


def test_nrm2_32_mixed_bands_and_specials():
    rng = np.random.default_rng(7)
    # entries above NU, inside [MU, NU] and below MU together: the naive sum
    # of squares overflows, the banded one does not
    x = rng.standard_normal(32)
    x[:5] *= 1e200
    x[5:10] *= 1e-200
    with np.errstate(over="ignore"):
        assert not np.isfinite(np.sum(x * x))
    ref = exact_norm(x)
    assert abs(o.nrm2_32(x) - ref) <= 4e-16 * ref
    y = rng.standard_normal(32) * 1e-170
    y[3] = 1.0
    assert abs(o.nrm2_32(y) - exact_norm(y)) <= 4e-16 * exact_norm(y)
    assert o.nrm2_32(np.zeros(32)) == 0.0
    z = np.zeros(32)
    z[17] = -3.5e-310          # a single subnormal
    assert o.nrm2_32(z) == 3.5e-310


@pytest.mark.parametrize("scale", [1.0, 1e-150, 1e150])
def test_cosa_against_exact(scale):
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.standard_normal(32) * scale
        y = (rng.standard_normal(32) + 0.3 * x / scale) * scale * rng.uniform(0.1, 10)
        ap, aq = o.nrm2_32(x), o.nrm2_32(y)
        c = o.cosa(ap, aq, x, y)
        dot = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y))
        ref = float(dot / (Fraction(exact_norm(x)) * Fraction(exact_norm(y))))
        assert abs(c - ref) <= 1e-14, (c, ref)
        assert abs(o.cosa(aq, ap, y, x) - c) <= 1e-15


def test_givens_defining_equations():
    rng = np.random.default_rng(5)
    for f, gg in list(zip(rng.standard_normal(200), rng.standard_normal(200) * 10.0 ** rng.uniform(-8, 8, 200))) \
            + [(0.0, 0.0), (3.0, 0.0), (-3.0, 0.0), (0.0, 2.0), (0.0, -2.0), (1e300, 1e300), (1e-300, 3e-300)]:
        c, s, r = o.givens(f, gg)
        assert r >= 0.0
        assert abs(r - math.hypot(f, gg)) <= 4e-16 * r
        if r > 0:
            # (the products and sums of the check itself round too)
            assert abs(c * f + s * gg - r) <= 1e-15 * r
            assert abs(-s * f + c * gg) <= 1e-15 * r
            assert abs(c * c + s * s - 1.0) <= 1e-15
    assert o.givens(3.0, 4.0) == (0.6, 0.8, 5.0)


def test_qr32_and_peeloff():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((32, 32)) * 10.0 ** rng.uniform(-5, 5, 32)
    R = o.qr32(A)
    assert np.all(np.tril(R, -1) == 0.0)
    d = np.sqrt(np.diag(A.T @ A))
    assert (np.abs(R.T @ R - A.T @ A) / np.outer(d, d)).max() < 1e-14
    Rn = np.linalg.qr(A, mode="r")
    assert (np.abs(np.abs(R) - np.abs(Rn)) / d).max() < 1e-13
    # merging two triangular factors
    B = rng.standard_normal((32, 32))
    R0, R1 = o.qr32(A), o.qr32(B)
    M, Z = o.peeloff(R0, R1)
    assert np.all(np.tril(M, -1) == 0.0) and np.all(Z == 0.0)
    assert np.all(np.diag(M) >= 0.0)
    S = R0.T @ R0 + R1.T @ R1
    dd = np.sqrt(np.diag(S))
    assert (np.abs(M.T @ M - S) / np.outer(dd, dd)).max() < 1e-14


@pytest.mark.parametrize("m,lo,hi", [(64, -1, 1), (1024, -8, 8), (512, -150, 150)])
def test_factorize_qr(m, lo, hi):
    rng = np.random.default_rng(m)
    G = rng.standard_normal((m, 32)) * 10.0 ** rng.uniform(lo, hi, 32)
    R = o.factorize_qr(G)
    assert np.all(np.tril(R, -1) == 0.0) and np.all(np.isfinite(R))
    Rn = np.linalg.qr(G, mode="r")
    d = np.sqrt(np.einsum("ij,ij->j", G / np.abs(G).max(0), G / np.abs(G).max(0))) * np.abs(G).max(0)
    # |R| is unique; compare column-scaled
    assert (np.abs(np.abs(R) - np.abs(Rn)) / d).max() < 1e-13
    if hi <= 8:
        assert (np.abs(R.T @ R - G.T @ G) / np.outer(d, d)).max() < 1e-13


def _strats(n):
    s0 = g.strat_generate("cycwor", 32).astype(np.uint8)
    s1 = g.strat_generate("cycwor", n // 16).astype(np.uint16)
    return s0, s1


@pytest.mark.parametrize("fact", [o.FACT_QR, o.FACT_QR_DRMAC])
@pytest.mark.parametrize("alg,kind", [(9, "svd"), (11, "svd"), (8, "svd"), (1, "svd"), (11, "hyp"), (3, "hyp")])
def test_qr_paths_converge_to_the_spectrum(fact, alg, kind):
    m, n = 256, 128
    rng = np.random.default_rng(alg + 10 * fact)
    s = 10.0 ** (-10.0 * rng.uniform(size=n))
    if kind == "svd":
        G, nplus = svd_input(m, n, s, seed=alg), n
    else:
        G, nplus = graded_input(m, n, seed=alg), n // 3
    s0, s1 = _strats(n)
    old = o.set_factorization(fact)
    try:
        r = o.hypJacL2(alg, G, nplus, s0, s1, nthreads=4)
    finally:
        o.set_factorization(old)
    rc = o.hypJacL2(alg, G, nplus, s0, s1, nthreads=4)
    assert r["ret"] == 0 and r["rot_b_trace"][-1] == 0
    sig = np.sort(np.sqrt(np.abs(r["D"])))[::-1]
    sigc = np.sort(np.sqrt(np.abs(rc["D"])))[::-1]
    assert np.abs(sig - sigc).max() <= 1e-13 * sig[0]
    if kind == "svd":
        ex = np.sort(s)[::-1]
        assert np.abs(sig - ex).max() <= 1e-13 * ex[0]
    if r["V"] is not None:
        J = np.r_[np.ones(nplus), -np.ones(n - nplus)]
        V = r["V"]
        assert np.abs(V.T @ (J[:, None] * V) - np.diag(J)).max() < 1e-11 * max(1.0, np.abs(V).max() ** 2)


def test_drmac_differs_only_on_sorted_definite_blocks():
    # Drmac's level-0 replaces the reference's only for sorted definite
    # blocks: unsorted runs of QR and QR+Drmac are the same bits
    G = graded_input(256, 64, seed=4)
    s0, s1 = _strats(64)
    outs = []
    for fact in (o.FACT_QR, o.FACT_QR_DRMAC):
        old = o.set_factorization(fact)
        try:
            outs.append(o.hypJacL2(8, G, 64, s0, s1, nthreads=2))
        finally:
            o.set_factorization(old)
    np.testing.assert_array_equal(outs[0]["G"], outs[1]["G"])
    np.testing.assert_array_equal(outs[0]["D"], outs[1]["D"])


@pytest.mark.parametrize("fact", [o.FACT_CHOLESKY, o.FACT_QR, o.FACT_QR_DRMAC])
@pytest.mark.parametrize("alg", [1, 9])
@pytest.mark.parametrize("sign", [+1, -1])
def test_sorted_runs_order_every_block(fact, alg, sign):
    # the sorted level-0 leaves every 16-column block ordered: larger norms
    # first for J = I (posd), smaller first for J = -I (negd); a swapped sort
    # rule in any of the three level-0 variants breaks this
    m, n = 256, 128
    rng = np.random.default_rng(3)
    s = 10.0 ** (-10.0 * rng.uniform(size=n))
    G = svd_input(m, n, s, seed=5)
    s0, s1 = _strats(n)
    old = o.set_factorization(fact)
    try:
        r = o.hypJacL2(alg, G, n if sign > 0 else 0, s0, s1, nthreads=4)
    finally:
        o.set_factorization(old)
    sig = np.sqrt(np.abs(r["D"])).reshape(-1, 16)
    assert np.all(np.sign(r["D"]) == sign)
    steps = np.diff(sig, axis=1)
    assert np.all(steps <= 0.0) if sign > 0 else np.all(steps >= 0.0)
    assert np.abs(np.sort(sig.ravel()) - np.sort(s)).max() <= 1e-13 * s.max()
