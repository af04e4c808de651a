"""CPU FP64 oracle of the blocked one-sided (hyperbolic) Jacobi SVD.

TEST INFRASTRUCTURE ONLY: the checker of the gfx950 library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.  It
wraps oracle/libhjoracle.so (hjoracle.c: a C restatement of the reference's
algorithm, 1/HypJacL2.cu and share/device_code_*.hpp, in the device's exact
operation order).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(_HERE, "libhjoracle.so")
REF_STRAT_SO = os.path.join(_HERE, "_ref", "strat.so")

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(SO):
            raise RuntimeError("oracle/libhjoracle.so missing: run `python -m gpujachx_amd.build`")
        L = ctypes.CDLL(SO)
        d, u, i = ctypes.c_double, ctypes.c_uint, ctypes.c_int
        pd, pu, pi = ctypes.POINTER(d), ctypes.POINTER(u), ctypes.POINTER(i)
        pull = ctypes.POINTER(ctypes.c_ulonglong)
        L.hjo_rsqrt.argtypes = [d]
        L.hjo_rsqrt.restype = d
        L.hjo_rsqrt_n.argtypes = [pd, pd, ctypes.c_long]
        L.hjo_rsqrt_n.restype = None
        L.hjo_rot_t.argtypes = [d, d, d, pd, pd]
        L.hjo_rot_t.restype = i
        L.hjo_rot_h.argtypes = [d, d, d, pd, pd]
        L.hjo_rot_h.restype = i
        L.hjo_sum32.argtypes = [pd]
        L.hjo_sum32.restype = d
        L.hjo_gram_nchunk.argtypes = [u, u]
        L.hjo_gram_nchunk.restype = u
        L.hjo_gram_partial.argtypes = [pd, pd, u, u, u, pd]
        L.hjo_gram_partial.restype = None
        L.hjo_pair.argtypes = [pd, u, i, i, i, u, ctypes.c_void_p, u, u, pd, pu, pi, pd]
        L.hjo_pair.restype = u
        L.hjo_update.argtypes = [pd, pd, u, u, pd]
        L.hjo_update.restype = None
        L.hjo_initD.argtypes = [pd, u, u, u, u, pd]
        L.hjo_initD.restype = None
        L.hjo_hypJacL2.argtypes = [u, u, u, u, pd, u, pd, u, pd, pu, pull, pull,
                                   ctypes.c_void_p, u, ctypes.c_void_p, u, i, ctypes.c_long, pu, pu]
        L.hjo_hypJacL2.restype = i
        L.hjo_hypJacL2_w.argtypes = [u, u, u, u, pd, u, pd, u, pd, pu, pull, pull, u,
                                     ctypes.c_void_p, u, ctypes.c_void_p, u, u, i, pu]
        L.hjo_hypJacL2_w.restype = i
        L.hjo_gram_partial_w.argtypes = [pd, u, u, u, u, u, u, pd]
        L.hjo_gram_partial_w.restype = None
        L.hjo_pair_w.argtypes = [pd, u, u, i, i, i, u, ctypes.c_void_p, u, u, pd, pu, pi]
        L.hjo_pair_w.restype = u
        L.hjo_update_w.argtypes = [pd, u, u, u, u, u, pd]
        L.hjo_update_w.restype = None
        L.hjo_set_max_sweeps.argtypes = [u]
        L.hjo_set_max_sweeps.restype = u
        L.hjo_set_single_wave_gram.argtypes = [i]
        L.hjo_set_single_wave_gram.restype = i
        L.hjo_set_factorization.argtypes = [i]
        L.hjo_set_factorization.restype = i
        L.hjo_nrm2_32.argtypes = [pd]
        L.hjo_nrm2_32.restype = d
        L.hjo_cosa.argtypes = [d, d, pd, pd]
        L.hjo_cosa.restype = d
        L.hjo_givens.argtypes = [d, d, pd, pd, pd]
        L.hjo_givens.restype = None
        L.hjo_qr32.argtypes = [pd]
        L.hjo_qr32.restype = None
        L.hjo_peeloff.argtypes = [pd, pd]
        L.hjo_peeloff.restype = None
        L.hjo_factorize_qr.argtypes = [pd, pd, u, u, pd]
        L.hjo_factorize_qr.restype = None
        _L = L
    return _L


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def rsqrt(x):
    return lib().hjo_rsqrt(float(x))


def rsqrt_n(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().hjo_rsqrt_n(_pd(x), _pd(y), x.size)
    return y


def rot_t(apq, dp, dq):
    c, t = ctypes.c_double(), ctypes.c_double()
    b = lib().hjo_rot_t(apq, dp, dq, ctypes.byref(c), ctypes.byref(t))
    return c.value, t.value, b


def rot_h(apq, dp, dq):
    c, t = ctypes.c_double(), ctypes.c_double()
    b = lib().hjo_rot_h(apq, dp, dq, ctypes.byref(c), ctypes.byref(t))
    return c.value, t.value, b


def sum32(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    assert v.size == 32
    return lib().hjo_sum32(_pd(v))


def gram_nchunk(nrow, npairs):
    return lib().hjo_gram_nchunk(nrow, npairs)


def hypJacL2(routine, G, nplus, strat0, strat1, ldhG=None, nthreads=0, max_steps=0):
    """Mirror of 1/HypJacL2.cu:13-202 on numpy data (G: nrow x ncol).
    strat0: uint8 [S0,16,2]; strat1: uint16 [S1,n1/2,2].
    Returns dict(ret, G(=U), V, D, sweeps, rot_s, rot_b, rot_s_trace, rot_b_trace)."""
    G = np.asarray(G, dtype=np.float64)
    nrow, ncol = G.shape
    ld = nrow if ldhG is None else ldhG
    hG = np.zeros((ncol, ld), dtype=np.float64)
    hG[:, :nrow] = G.T
    full = bool(routine & 2)
    hV = np.zeros((ncol, ncol), dtype=np.float64) if full else None
    hD = np.zeros(ncol, dtype=np.float64)
    s0 = np.ascontiguousarray(strat0, dtype=np.uint8)
    s1 = np.ascontiguousarray(strat1, dtype=np.uint16)
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    ts = (ctypes.c_uint * 256)()
    tb = (ctypes.c_uint * 256)()
    ret = lib().hjo_hypJacL2(routine, nrow, ncol, nplus, _pd(hG), ld, _pd(hV), ncol if full else 0,
                             _pd(hD), ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb),
                             s0.ctypes.data, s0.shape[0], s1.ctypes.data, s1.shape[0],
                             int(nthreads), int(max_steps), ts, tb)
    k = sw.value
    return dict(ret=ret, G=hG[:, :nrow].T.copy(), V=None if hV is None else hV.T.copy(), D=hD,
                sweeps=k, rot_s=gs.value, rot_b=gb.value,
                rot_s_trace=list(ts[:k]), rot_b_trace=list(tb[:k]))


def hypJacL2_w(routine, G, nplus, width, strat0, strat1, max_sweeps=30):
    """The method with blocks of `width` (16 or 32) columns: strat0 uint8
    [S0, width, 2] of order 2 width, strat1 uint16 [S1, ncol/(2 width), 2] of
    order ncol / width.  Returns the dict of hypJacL2 (rot_b_trace too)."""
    G = np.asarray(G, dtype=np.float64)
    nrow, ncol = G.shape
    hG = np.ascontiguousarray(G.T)
    full = bool(routine & 2)
    hV = np.zeros((ncol, ncol), dtype=np.float64) if full else None
    hD = np.zeros(ncol, dtype=np.float64)
    s0 = np.ascontiguousarray(strat0, dtype=np.uint8)
    s1 = np.ascontiguousarray(strat1, dtype=np.uint16)
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    tb = (ctypes.c_uint * 256)()
    ret = lib().hjo_hypJacL2_w(routine, nrow, ncol, nplus, _pd(hG), nrow, _pd(hV),
                               ncol if full else 0, _pd(hD), ctypes.byref(sw), ctypes.byref(gs),
                               ctypes.byref(gb), int(width), s0.ctypes.data, s0.shape[0],
                               s1.ctypes.data, s1.shape[0], int(max_sweeps), 0, tb)
    k = sw.value
    return dict(ret=ret, G=hG.T.copy(), V=None if hV is None else hV.T.copy(), D=hD,
                sweeps=k, rot_s=gs.value, rot_b=gb.value, rot_b_trace=list(tb[:k]))


def gram_partial_w(G, width, a, b, r0, nr):
    """Gram partial (tiles of the pair's lower triangle, MFMA layout) of
    rows [r0, r0+nr) of the pair of width-column blocks a, b of G."""
    hG = np.ascontiguousarray(np.asarray(G, dtype=np.float64).T)
    nt = 2 * width // 16
    out = np.zeros(nt * (nt + 1) // 2 * 256)
    lib().hjo_gram_partial_w(_pd(hG), G.shape[0], int(width), int(a), int(b), int(r0), int(nr),
                             _pd(out))
    return out


def tiles_to_lower(part, width):
    """The pair's Gram (lower triangle, zeros above) from one partial."""
    n2 = 2 * width
    nt = n2 // 16
    A = np.zeros((n2, n2))
    tt = 0
    for ti in range(nt):
        for tj in range(ti + 1):
            t = part[tt * 256:(tt + 1) * 256].reshape(64, 4)
            for l in range(64):
                for r in range(4):
                    A[16 * ti + 4 * r + l // 16, 16 * tj + l % 16] = t[l, r]
            tt += 1
    return A


def pair_w(partials, width, sort, hyp, definite, npos, strat0, nswp=1):
    """One pair of width-column blocks from its Gram partials; (W, s, b, upd)."""
    p = np.ascontiguousarray(partials, dtype=np.float64).reshape(-1)
    n2 = 2 * width
    nt = n2 // 16
    nchunk = p.size // (nt * (nt + 1) // 2 * 256)
    s0 = np.ascontiguousarray(strat0, dtype=np.uint8)
    W = np.zeros(n2 * n2)
    b = ctypes.c_uint(0)
    upd = ctypes.c_int(0)
    s = lib().hjo_pair_w(_pd(p), nchunk, int(width), int(sort), int(hyp), int(definite), int(npos),
                         s0.ctypes.data, s0.shape[0], nswp, _pd(W), ctypes.byref(b),
                         ctypes.byref(upd))
    return W.reshape(n2, n2).T.copy(), s, b.value, upd.value


def update_w(G, width, a, b, Wm):
    """G with the pair (a, b) of width-column blocks multiplied by Wm."""
    hG = np.ascontiguousarray(np.asarray(G, dtype=np.float64).T)
    Wc = np.ascontiguousarray(np.asarray(Wm, dtype=np.float64).T)
    lib().hjo_update_w(_pd(hG), G.shape[0], int(width), int(a), int(b), G.shape[0], _pd(Wc))
    return hG.T.copy()


def set_single_wave_gram(on):
    """Gram order of the batched launch (one wave over all rows)."""
    return lib().hjo_set_single_wave_gram(1 if on else 0)


def set_max_sweeps(n):
    return lib().hjo_set_max_sweeps(int(n))


def pair(partials, sort, hyp, definite, npos, strat0, nswp=1):
    """One block pair from its Gram partials; returns (W, R, s, b, upd)."""
    p = np.ascontiguousarray(partials, dtype=np.float64).reshape(-1)
    nchunk = p.size // 768
    s0 = np.ascontiguousarray(strat0, dtype=np.uint8)
    W = np.zeros(1024)
    R = np.zeros(1024)
    b = ctypes.c_uint(0)
    upd = ctypes.c_int(0)
    s = lib().hjo_pair(_pd(p), nchunk, int(sort), int(hyp), int(definite), int(npos),
                       s0.ctypes.data, s0.shape[0], nswp, _pd(W), ctypes.byref(b), ctypes.byref(upd),
                       _pd(R))
    return W.reshape(32, 32).T.copy(), R.reshape(32, 32).T.copy(), s, b.value, upd.value


# factorization of the block pairs (the reference's build options)
FACT_CHOLESKY, FACT_QR, FACT_QR_DRMAC = 0, 1, 2


def set_factorization(f):
    """Cholesky of the Gram (0), QR by 32-row panels (1), QR with Drmac's
    sorted level-0 (2); returns the previous setting."""
    return lib().hjo_set_factorization(int(f))


def nrm2_32(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.shape == (32,)
    return lib().hjo_nrm2_32(_pd(x))


def cosa(ap, aq, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().hjo_cosa(ap, aq, _pd(x), _pd(y))


def givens(f, g):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().hjo_givens(f, g, ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def qr32(A):
    """R of a 32x32 matrix by dQR32 (returns a new array)."""
    F = np.asfortranarray(np.array(A, dtype=np.float64))
    buf = F.T.ravel().copy()          # column-major
    lib().hjo_qr32(_pd(buf))
    return buf.reshape(32, 32).T.copy()


def peeloff(R0, R1):
    """R of [R0; R1] for upper triangular R0, R1 (dPeelOff); returns (R, R1')."""
    b0 = np.asarray(R0, dtype=np.float64).T.ravel().copy()
    b1 = np.asarray(R1, dtype=np.float64).T.ravel().copy()
    lib().hjo_peeloff(_pd(b0), _pd(b1))
    return b0.reshape(32, 32).T.copy(), b1.reshape(32, 32).T.copy()


def factorize_qr(G):
    """R (32x32) of an nrow x 32 column pair by 32-row panels."""
    G = np.asarray(G, dtype=np.float64)
    nrow = G.shape[0]
    buf = G.T.ravel().copy()          # column-major, ld = nrow
    R = np.zeros(1024)
    L = lib()
    base = buf.ctypes.data
    L.hjo_factorize_qr(ctypes.cast(base, ctypes.POINTER(ctypes.c_double)),
                       ctypes.cast(base + 16 * nrow * 8, ctypes.POINTER(ctypes.c_double)),
                       nrow, nrow, _pd(R))
    return R.reshape(32, 32).T.copy()
