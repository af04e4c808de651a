"""Accuracy measures of (hyperbolic) singular values, per value and
normwise, shared by the benchmarks and the tests (cf. the per-eigenvalue
RELERR of var/pp80.f90:768-841)."""
import numpy as np


def per_sigma(a, b):
    """Per-value agreement of two spectra, b the reference: 1-D arrays of
    sigma, or of D = j sigma^2 from a hyperbolic SVD (the positive and the
    negative parts are compared separately, by absolute value).  After
    sorting: relative differences |a_i - b_i| / |b_i| (max, median, the
    fraction within 1e-12, the reference value at the worst one, the largest
    reference value whose difference exceeds 1e-12) and the normwise
    max |a_i - b_i| / max |b|."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    pa, pb = np.sort(a[a >= 0]), np.sort(b[b >= 0])
    na, nb = np.sort(-a[a < 0]), np.sort(-b[b < 0])
    if pa.size != pb.size or na.size != nb.size:
        return {"inertia_mismatch": [int(pa.size), int(na.size), int(pb.size), int(nb.size)]}
    va, vb = np.concatenate([pa, na]), np.concatenate([pb, nb])
    rel = np.abs(va - vb) / vb
    worst = int(np.argmax(rel))
    bad = vb[rel > 1e-12]
    return {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
            "frac_rel_le_1e-12": float(np.mean(rel <= 1e-12)),
            "worst_at_value": float(vb[worst]),
            "largest_value_above_1e-12": float(bad.max()) if bad.size else None,
            "normwise": float(np.max(np.abs(va - vb)) / np.max(vb))}


def rel_diffs(a, b):
    """The sorted per-value relative differences of per_sigma, as an array
    aligned with the sorted reference values (positive part, then negative)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    va = np.concatenate([np.sort(a[a >= 0]), np.sort(-a[a < 0])])
    vb = np.concatenate([np.sort(b[b >= 0]), np.sort(-b[b < 0])])
    if va.size != vb.size:
        raise ValueError("spectra of different inertia")
    return np.abs(va - vb) / vb, vb


def exact_err(sig, ex):
    """sigma against a constructed spectrum ex (both 1-D, any order)."""
    sig_s, ex_s = np.sort(sig)[::-1], np.sort(ex)[::-1]
    rel = np.abs(sig_s - ex_s) / ex_s
    return {"max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
            "normwise": float(np.max(np.abs(sig_s - ex_s)) / ex_s[0]),
            "frac_rel_le_1e-12": float(np.mean(rel <= 1e-12))}
