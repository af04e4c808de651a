"""gpujachx_amd -- blocked one-sided (hyperbolic) Jacobi SVD on MI355X (gfx950).

The product is the native library ``lib/libhypjac.so`` (HIP kernels + C ABI,
see ``include/hypjac.h``).  This module is a thin ctypes mirror of the
reference's interface (``hypJacL2``, ``init_strats``, ``configureGPU``) with
numpy buffers, plus access to the library's additions.  It never falls back to
a CPU implementation: if the library is missing, importing raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# HJ_LIBHYPJAC: path of another build of the same library
LIB_PATH = os.environ.get("HJ_LIBHYPJAC") or os.path.join(LIB_DIR, "libhypjac.so")
STRAT_SO = os.path.join(LIB_DIR, "strat.so")

# ALG bits (share/HypJacL.hpp:88-110)
BLK_ORI = 8
ACCUMV_2 = 4
FULL_SVD = 2
CDSORT_1 = 1
NSWEEP = 30

FAMILIES = ("mmstep", "BrentL", "colcyc", "cycloc", "rowcyc", "cycwor")

# forms of the method (include/hypjac.h)
FORM_PLAIN = 0
FORM_LEARNED_SKIP = 1
FORM_SPLIT_GRAM = 2

_lib = None


class BatchDesc(ctypes.Structure):
    """hj_batch_desc: one problem of a batch (device pointers)."""
    _fields_ = [("routine", ctypes.c_uint), ("nrow", ctypes.c_uint), ("ncol", ctypes.c_uint),
                ("nplus", ctypes.c_uint), ("ldG", ctypes.c_uint), ("ldV", ctypes.c_uint),
                ("G", ctypes.c_void_p), ("V", ctypes.c_void_p), ("D", ctypes.c_void_p)]


class BatchCfg(ctypes.Structure):
    """hj_batch_cfg: launch shape and schedule of a batch."""
    _fields_ = [("waves", ctypes.c_uint), ("wg_per_cu", ctypes.c_uint), ("strategy", ctypes.c_uint),
                ("max_sweeps", ctypes.c_uint), ("longest_first", ctypes.c_uint)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BATCH_STRAT = {"cycwor": 0, "BrentL": 1, "mmstep": 2}


def lib():
    """Load libhypjac.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libhypjac.so is missing (%s): build it with `python -m gpujachx_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    u, i, d = ctypes.c_uint, ctypes.c_int, ctypes.c_double
    pd, pu = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint)
    pull = ctypes.POINTER(ctypes.c_ulonglong)
    vp, cp = ctypes.c_void_p, ctypes.c_char_p
    L.hypJacL2.argtypes = [u, u, u, u, pd, u, pd, u, pd, pu, pull, pull, pd]
    L.hypJacL2.restype = i
    L.hypJacL2_dev.argtypes = [u, u, u, u, vp, u, vp, u, vp, pu, pull, pull, pd, vp]
    L.hypJacL2_dev.restype = i
    L.hypJacL2_dev_noalloc.argtypes = [u, u, u, u, vp, u, vp, u, vp, pu, pull, pull, pd, vp]
    L.hypJacL2_dev_noalloc.restype = i
    L.init_strats.argtypes = [cp, cp, u, cp, u]
    L.init_strats.restype = None
    L.hj_init_strats_builtin.argtypes = [cp, u, cp, u]
    L.hj_init_strats_builtin.restype = None
    L.hj_strat_generate.argtypes = [cp, u, ctypes.POINTER(ctypes.c_ushort), pu]
    L.hj_strat_generate.restype = i
    L.configureGPU.argtypes = [i]
    L.configureGPU.restype = i
    L.hj_last_sweeps.argtypes = [pu, pu, pu, pd]
    L.hj_last_sweeps.restype = None
    L.hj_set_profile.argtypes = [i]
    L.hj_set_profile.restype = None
    L.hj_kernel_ms.argtypes = [i, pull]
    L.hj_kernel_ms.restype = d
    L.hj_gram_nchunk.argtypes = [u, u]
    L.hj_gram_nchunk.restype = u
    L.hj_set_max_sweeps.argtypes = [u]
    L.hj_set_max_sweeps.restype = u
    L.hj_set_step_limit.argtypes = [ctypes.c_long]
    L.hj_set_step_limit.restype = None
    L.hj_set_factorization.argtypes = [i]
    L.hj_set_factorization.restype = i
    L.hj_set_staged.argtypes = [i]
    L.hj_set_staged.restype = i
    L.hj_last_staged.argtypes = []
    L.hj_last_staged.restype = i
    L.hj_dbg_scalar.argtypes = [i, i, pd, pd, pd, pd, pd, ctypes.POINTER(ctypes.c_int)]
    L.hj_dbg_scalar.restype = i
    L.hj_version.argtypes = []
    L.hj_version.restype = cp
    pf = ctypes.POINTER(ctypes.c_float)
    L.hj_set_form.argtypes = [u]
    L.hj_set_form.restype = None
    L.hj_set_skip_model.argtypes = [pf, u, ctypes.c_float, u]
    L.hj_set_skip_model.restype = i
    L.hj_mlp_nparam.restype = u
    L.hj_mlp_nfeat.restype = u
    L.hj_visit_record_size.restype = u
    L.hj_set_visit_trace.argtypes = [vp, u]
    L.hj_set_visit_trace.restype = None
    L.hj_last_skips.argtypes = [pu, pull, pull]
    L.hj_last_skips.restype = None
    L.hj_skip_features.argtypes = [vp, u, ctypes.POINTER(ctypes.c_ushort), u, u, u, ctypes.c_float, pf]
    L.hj_skip_features.restype = i
    L.hj_mlp_grad.argtypes = [pf, pf, u, pf, ctypes.c_float, pf, pf]
    L.hj_mlp_grad.restype = i
    L.hj_mlp_predict.argtypes = [pf, u, pf, pf]
    L.hj_mlp_predict.restype = i
    L.hj_mlp_train.argtypes = [pf, pf, u, u, u, ctypes.c_float, ctypes.c_float, u, pf, pf]
    L.hj_mlp_train.restype = i
    L.hj_set_skip_min_frac.argtypes = [ctypes.c_float]
    L.hj_set_skip_min_frac.restype = None
    L.hj_set_skip_floor_mixed.argtypes = [ctypes.c_float]
    L.hj_set_skip_floor_mixed.restype = None
    L.hj_set_skip_min_propose.argtypes = [ctypes.c_float]
    L.hj_set_skip_min_propose.restype = None
    L.hj_set_skip_min_visits.argtypes = [ctypes.c_uint]
    L.hj_set_skip_min_visits.restype = None
    L.hj_set_split_tau.argtypes = [ctypes.c_float]
    L.hj_set_split_tau.restype = None
    L.hj_set_fused_split.argtypes = [ctypes.c_int]
    L.hj_set_fused_split.restype = ctypes.c_int
    L.hj_set_block_width.argtypes = [u]
    L.hj_set_block_width.restype = u
    L.hj_last_splits.argtypes = [pu, pull]
    L.hj_last_splits.restype = None
    L.hj_dbg_gram.argtypes = [i, u, pd, pd]
    L.hj_dbg_gram.restype = i
    pdesc, pcfg = ctypes.POINTER(BatchDesc), ctypes.POINTER(BatchCfg)
    L.hj_batch_default_cfg.argtypes = [pcfg]
    L.hj_batch_default_cfg.restype = None
    L.hj_batch_plan.argtypes = [pdesc, u, pcfg, ctypes.POINTER(i)]
    L.hj_batch_plan.restype = vp
    L.hj_batch_run.argtypes = [vp, vp]
    L.hj_batch_run.restype = i
    L.hj_batch_results.argtypes = [vp, pu, pull, pull, pd, pd]
    L.hj_batch_results.restype = i
    L.hj_batch_free.argtypes = [vp]
    L.hj_batch_free.restype = None
    L.hj_batch_autotune.argtypes = [pdesc, u, u, pcfg, pd]
    L.hj_batch_autotune.restype = i
    _lib = L
    return L


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def strat_generate(fam, n):
    """Built-in table of family `fam`, order n: uint16 array [steps, n/2, 2]."""
    L = lib()
    steps = ctypes.c_uint(0)
    if L.hj_strat_generate(fam.encode(), n, None, ctypes.byref(steps)) != 0:
        raise ValueError("no construction for %s %d" % (fam, n))
    out = np.zeros((steps.value, n // 2, 2), dtype=np.uint16)
    L.hj_strat_generate(fam.encode(), n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)),
                        ctypes.byref(steps))
    return out


def init_strats(sdy, snp0, n0, snp1, n1):
    """share/HypJacL.hpp:206 -- load tables from a strat.so (fail-stop)."""
    lib().init_strats(sdy.encode(), snp0.encode(), n0, snp1.encode(), n1)


def init_strats_builtin(snp0, n0, snp1, n1):
    """Same tables from the library's own generator (no strat.so)."""
    lib().hj_init_strats_builtin(snp0.encode(), n0, snp1.encode(), n1)


def current_strats():
    """(strat0 [S0,16,2] uint8, strat1 [S1,P1,2] uint16) currently loaded."""
    L = lib()
    s0 = ctypes.c_uint.in_dll(L, "STRAT0_STEPS").value
    p0 = ctypes.c_uint.in_dll(L, "STRAT0_PAIRS").value
    s1 = ctypes.c_uint.in_dll(L, "STRAT1_STEPS").value
    p1 = ctypes.c_uint.in_dll(L, "STRAT1_PAIRS").value
    t0 = np.ctypeslib.as_array((ctypes.c_ubyte * (32 * 16 * 2)).in_dll(L, "strat0")).reshape(32, 16, 2)
    t1 = np.ctypeslib.as_array((ctypes.c_ushort * (1024 * 512 * 2)).in_dll(L, "strat1")).reshape(1024, 512, 2)
    return t0[:s0, :p0].copy(), t1[:s1, :p1].copy()


def strat_ids():
    L = lib()
    return ctypes.c_uint.in_dll(L, "STRAT0").value, ctypes.c_uint.in_dll(L, "STRAT1").value


def configureGPU(dev=0):
    return lib().configureGPU(dev)


class Result:
    """Outputs of one hypJacL2 call."""

    def __init__(self, ret, G, V, D, sweeps, rot_s, rot_b, timing):
        self.ret, self.G, self.V, self.D = ret, G, V, D
        self.sweeps, self.rot_s, self.rot_b, self.timing = sweeps, rot_s, rot_b, timing

    @property
    def sigma(self):
        return np.sqrt(np.abs(self.D))


def hypJacL2(routine, G, nplus=None, ldhG=None, want_v=None):
    """1/HypJacL2.hpp:6-27 on numpy data.  G (nrow x ncol, float64) is copied
    into a Fortran-ordered buffer with leading dimension ldhG and overwritten
    by U; returns Result (ret < 0 on an argument error, like the reference)."""
    L = lib()
    G = np.asarray(G, dtype=np.float64)
    nrow, ncol = G.shape
    nplus = ncol if nplus is None else nplus
    ldhG = nrow if ldhG is None else ldhG
    hG = np.zeros((ncol, ldhG), dtype=np.float64)
    hG[:, :nrow] = G.T
    full = bool(routine & FULL_SVD) if want_v is None else want_v
    hV = np.zeros((ncol, ncol), dtype=np.float64) if full else None
    hD = np.zeros(ncol, dtype=np.float64)
    sw = ctypes.c_uint(0)
    gs, gb = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    tm = np.zeros(4, dtype=np.float64)
    ret = L.hypJacL2(routine, nrow, ncol, nplus, _pd(hG), ldhG, _pd(hV), ncol if full else 0,
                     _pd(hD), ctypes.byref(sw), ctypes.byref(gs), ctypes.byref(gb), _pd(tm))
    return Result(ret, hG[:, :nrow].T.copy(), None if hV is None else hV.T.copy(), hD,
                  sw.value, gs.value, gb.value, tm)


def last_sweeps():
    L = lib()
    n = ctypes.c_uint(0)
    s = (ctypes.c_uint * 256)()
    b = (ctypes.c_uint * 256)()
    t = (ctypes.c_double * 256)()
    L.hj_last_sweeps(ctypes.byref(n), s, b, t)
    k = n.value
    return list(s[:k]), list(b[:k]), list(t[:k])


def kernel_ms(kind):
    n = ctypes.c_ulonglong(0)
    ms = lib().hj_kernel_ms(kind, ctypes.byref(n))
    return ms, n.value


def gram_nchunk(nrow, npairs):
    return lib().hj_gram_nchunk(nrow, npairs)


def set_max_sweeps(n):
    """Level-2 sweep cap (the reference's HYPJAC_NSWEEP = 30 by default)."""
    return lib().hj_set_max_sweeps(int(n))


def set_step_limit(steps):
    lib().hj_set_step_limit(int(steps))


def set_form(form):
    lib().hj_set_form(int(form))


FACT_CHOLESKY, FACT_QR, FACT_QR_DRMAC = 0, 1, 2


def set_factorization(f):
    """Factorization of the block pairs: FACT_CHOLESKY (default), FACT_QR,
    FACT_QR_DRMAC (the reference's build options USE_QR / USE_DRMAC);
    returns the previous setting."""
    return lib().hj_set_factorization(int(f))


STAGED_OFF, STAGED_AUTO, STAGED_ALWAYS = 0, 1, 2


def set_staged(mode):
    """Staged sweeps (each step transition as two launches that overlap pair
    visits with half the block updates): STAGED_OFF, STAGED_AUTO (default:
    where a step's block updates move at least 4.8e8 bytes) or
    STAGED_ALWAYS; True means STAGED_AUTO.  Returns the previous mode."""
    return lib().hj_set_staged(int(mode))


def last_staged():
    """Whether the last hypJacL2 / hypJacL2_dev call ran staged sweeps."""
    return bool(lib().hj_last_staged())


def last_skips():
    """(per-sweep skipped visits, total skipped, total executed visits)."""
    n = len(last_sweeps()[0])
    arr = (ctypes.c_uint * 256)()
    sk, vi = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    lib().hj_last_skips(arr, ctypes.byref(sk), ctypes.byref(vi))
    return list(arr[:n]), sk.value, vi.value


def set_split_tau(tau):
    lib().hj_set_split_tau(float(tau))


def last_splits():
    """(per-sweep split-Gram visits, total) of the last call."""
    n = len(last_sweeps()[0])
    arr = (ctypes.c_uint * 256)()
    tot = ctypes.c_ulonglong(0)
    lib().hj_last_splits(arr, ctypes.byref(tot))
    return list(arr[:n]), tot.value


def dbg_gram(G, split):
    """32x32 Gram of an nrow x 32 matrix from the FP64 or the split kernel."""
    G = np.asfortranarray(G, dtype=np.float64)
    assert G.shape[1] == 32
    A = np.zeros((32, 32), dtype=np.float64, order="F")
    if lib().hj_dbg_gram(1 if split else 0, G.shape[0], _pd(G), _pd(A)):
        raise ValueError("hj_dbg_gram: nrow must be a positive multiple of 64")
    return np.ascontiguousarray(A)


def set_profile(on):
    lib().hj_set_profile(1 if on else 0)


def dbg_scalar(kind, a, b=None, c=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.size
    b = np.ascontiguousarray(a if b is None else b, dtype=np.float64)
    c = np.ascontiguousarray(a if c is None else c, dtype=np.float64)
    oc, ot = np.zeros(n), np.zeros(n)
    ob = np.zeros(n, dtype=np.int32)
    rc = lib().hj_dbg_scalar(kind, n, _pd(a), _pd(b), _pd(c), _pd(oc), _pd(ot),
                             ob.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc:
        raise RuntimeError("hj_dbg_scalar failed")
    return oc, ot, ob


class Batch:
    """A planned batch of independent (H)SVDs on device buffers (one
    persistent launch captured in a hipGraph).  `descs` is a list of
    BatchDesc; run() executes it on `stream` (a hipStream_t handle or None)."""

    def __init__(self, descs, cfg=None):
        L = lib()
        self.n = len(descs)
        self._arr = (BatchDesc * self.n)(*descs)
        err = ctypes.c_int(0)
        self.cfg = cfg if cfg is not None else default_batch_cfg()
        self._plan = L.hj_batch_plan(self._arr, self.n, ctypes.byref(self.cfg), ctypes.byref(err))
        if not self._plan:
            raise ValueError("hj_batch_plan failed with %d" % err.value)

    def run(self, stream=None):
        if lib().hj_batch_run(self._plan, stream):
            raise RuntimeError("hj_batch_run failed")

    def results(self):
        n = self.n
        sw = (ctypes.c_uint * n)()
        rs, rb = (ctypes.c_ulonglong * n)(), (ctypes.c_ulonglong * n)()
        t0, t1 = np.zeros(n), np.zeros(n)
        lib().hj_batch_results(self._plan, sw, rs, rb, _pd(t0), _pd(t1))
        return {"sweeps": np.array(sw[:]), "rot_s": np.array(rs[:]), "rot_b": np.array(rb[:]),
                "start_us": t0, "end_us": t1}

    def close(self):
        if self._plan:
            lib().hj_batch_free(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_batch_cfg():
    c = BatchCfg()
    lib().hj_batch_default_cfg(ctypes.byref(c))
    return c


def batch_autotune(descs, sample=256, max_sweeps=None):
    """(best BatchCfg, problems/s on the sample)."""
    arr = (BatchDesc * len(descs))(*descs)
    best = default_batch_cfg()
    if max_sweeps:
        best.max_sweeps = max_sweeps
    rate = ctypes.c_double(0.0)
    if lib().hj_batch_autotune(arr, len(descs), sample, ctypes.byref(best), ctypes.byref(rate)):
        raise RuntimeError("hj_batch_autotune failed")
    return best, rate.value
