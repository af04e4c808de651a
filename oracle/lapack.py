"""LAPACK dgesvj (one-sided Jacobi SVD, J = I) as an independent CPU
comparator, loaded with ctypes from the OpenBLAS that ships with scipy.
Test / CPU-baseline infrastructure only."""
import ctypes
import glob
import os

import numpy as np

_L = None


def _lib():
    global _L
    if _L is None:
        import scipy
        base = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        cands = sorted(glob.glob(os.path.join(base, "libscipy_openblas*.so")))
        if not cands:
            raise RuntimeError("no OpenBLAS with dgesvj found")
        L = ctypes.CDLL(cands[0])
        f = L.scipy_LAPACKE_dgesvj
        i, c, pd = ctypes.c_int, ctypes.c_char, ctypes.POINTER(ctypes.c_double)
        f.argtypes = [i, c, c, c, i, i, pd, i, pd, i, pd, i, pd]
        f.restype = i
        _L = L
    return _L


def dgesvj(G, want_v=True):
    """sigma (as LAPACK orders them), U, V of G (m x n, m >= n)."""
    A = np.asfortranarray(np.array(G, dtype=np.float64))
    m, n = A.shape
    sva = np.zeros(n)
    V = np.zeros((n, n), order="F") if want_v else np.zeros((1, 1), order="F")
    stat = np.zeros(6)
    pd = ctypes.POINTER(ctypes.c_double)
    info = _lib().scipy_LAPACKE_dgesvj(102, b"G", b"U", b"V" if want_v else b"N", m, n,
                                       A.ctypes.data_as(pd), m, sva.ctypes.data_as(pd),
                                       0, V.ctypes.data_as(pd), n if want_v else 1,
                                       stat.ctypes.data_as(pd))
    if info < 0:
        raise ValueError("dgesvj argument %d" % -info)
    # sva holds scaled sigmas: sigma = stat[0] * sva
    return stat[0] * sva, A, (V if want_v else None), info, stat
