"""Writes the golden fixtures of tests/golden/ (run from the repo root):

  rsqrt.json       correctly rounded 1/sqrt(x) from mpmath (200-bit) for
                   random and edge inputs -- pins the oracle's own rsqrt.
  rotations.json   exact-arithmetic cosine/tangent of the 2x2 trigonometric
                   and hyperbolic Jacobi rotations (mpmath, 200-bit) on the
                   branch of share/device_code_common_rotate.hpp:15-64 that
                   the double-precision formula takes, plus the branch flag.
  strat_fingerprints.json  sha256 (first 16 hex digits) of the reference's
                   strategy tables flattened [S][P][2] as uint16 LE, as
                   recorded from /root/reference/strat (families x orders).

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import re

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.prec = 200


def rsqrt_cases():
    rng = np.random.default_rng(20261202)
    xs = list(rng.uniform(0.5, 4.0, 3000)) + list(10.0 ** rng.uniform(-300, 300, 2000))
    xs += list(1.0 + rng.uniform(0, 1, 1000) ** 12)  # 1 + t^2 near 1
    xs += list(1.0 - rng.uniform(0, 1, 1000) ** 4)   # 1 - t^2
    xs += [1.0, 2.0, 4.0, 0.25, 0.5, 3.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
           np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), np.nextafter(4.0, 0.0)]
    out = []
    for x in xs:
        x = float(x)
        y = float(1 / mpmath.sqrt(mpmath.mpf(x)))  # mpf -> float rounds to nearest
        out.append([x.hex(), y.hex()])
    return out


def rot_cases():
    rng = np.random.default_rng(7)
    sq_heps, sq_2heps = 2.0 ** -26.5, 2.0 ** 27
    cases = []
    n = 1500
    apq = rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 1, n)
    dp = 10.0 ** rng.uniform(-2, 2, n)
    dq = dp * (1.0 + rng.standard_normal(n) * 10.0 ** rng.uniform(-12, 1, n))
    dq = np.abs(dq)
    for a, p, q in zip(apq, dp, dq):
        for kind in ("t", "h"):
            A, P, Q = mpmath.mpf(float(a)), mpmath.mpf(float(p)), mpmath.mpf(float(q))
            # branch as the double formula decides it (same ct2 the oracle computes)
            if kind == "t":
                ct2 = (float(q) - float(p)) / (2.0 * float(a))
                CT2 = abs(ct2)
                z = (Q - P) / (2 * A)
                if CT2 < sq_heps:
                    br = 0
                    t = mpmath.sign(z) / (abs(z) + 1)
                elif CT2 >= sq_2heps:
                    br = 1
                    t = mpmath.sign(z) / (2 * abs(z))
                else:
                    br = 2
                    t = mpmath.sign(z) / (abs(z) + mpmath.sqrt(1 + z * z))
                c = 1 / mpmath.sqrt(1 + t * t) if br != 1 else mpmath.mpf(1)
            else:
                ct2 = -((float(q) + float(p)) / (2.0 * float(a)))
                CT2 = abs(ct2)
                z = -(Q + P) / (2 * A)
                if CT2 >= sq_2heps:
                    br = 1
                    t = mpmath.sign(z) / (2 * abs(z))
                    c = mpmath.mpf(1)
                elif CT2 > 1.0:
                    br = 2
                    t = mpmath.sign(z) / (abs(z) + mpmath.sqrt(z * z - 1))
                    c = 1 / mpmath.sqrt(1 - t * t)
                else:
                    br = 3
                    t, c = mpmath.mpf(0.5), 2 / mpmath.sqrt(3)
            cases.append([kind, float(a).hex(), float(p).hex(), float(q).hex(), br,
                          float(c).hex(), float(t).hex()])
    return cases


def fingerprints():
    ref = "/root/reference/strat"
    out = {}
    if not os.path.isdir(ref):
        return None
    for fam in ("cycwor", "rowcyc", "colcyc", "cycloc", "BrentL", "mmstep"):
        for n in (16, 32, 128, 512):
            txt = open(os.path.join(ref, fam, "%05d.h" % n)).read()
            m = re.search(r"\[\s*(\d+)\]\[\s*(\d+)\]\[2\]", txt)
            S, P = int(m.group(1)), int(m.group(2))
            nums = np.array([int(v) for v in re.findall(r"\d+", txt[txt.index("{"):])], dtype="<u2")
            assert nums.size == S * P * 2
            out["%s%05d" % (fam, n)] = [S, P, hashlib.sha256(nums.tobytes()).hexdigest()[:16]]
    return out


def main():
    json.dump(rsqrt_cases(), open(os.path.join(HERE, "rsqrt.json"), "w"))
    json.dump(rot_cases(), open(os.path.join(HERE, "rotations.json"), "w"))
    fp = fingerprints()
    if fp is not None:
        json.dump(fp, open(os.path.join(HERE, "strat_fingerprints.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
