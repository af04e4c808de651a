"""The C-ABI library: loads without a GPU, exports every entry point that
include/hypjac.h declares, and validates arguments exactly like
1/HypJacL2.cu:39-77 (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpujachx_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hypjac.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def declared_data():
    txt = open(os.path.join(ROOT, "include", "hypjac.h")).read()
    return re.findall(r"extern\s+unsigned\s+(?:char|short)\s+(\w+)\[", txt) + \
        [v for line in re.findall(r"extern unsigned ([A-Z0-9_, ]+);", txt) for v in
         [x.strip() for x in line.split(",")]]


def test_library_loads_and_exports_every_symbol():
    L = g.lib()
    funcs = declared_functions()
    assert "hypJacL2" in funcs and "init_strats" in funcs and len(funcs) >= 28
    missing = [f for f in funcs if not hasattr(L, f)]
    assert not missing, missing
    for v in declared_data():
        ctypes.c_uint.in_dll(L, v)  # raises if absent
    assert g.lib().hj_version().startswith(b"gpujachx_amd")


def test_strat_plugin_symbols():
    L = ctypes.CDLL(g.STRAT_SO)
    for fam in g.FAMILIES:
        for n in (2, 32, 512):
            ctypes.c_ubyte.in_dll(L, "%s%05u" % (fam, n))
    ctypes.c_ubyte.in_dll(L, "cycwor01024")


def call(routine, nrow, ncol, nplus, G=True, ldhG=None, V=True, ldhV=None, D=True,
         sw=True, gs=True, gb=True):
    L = g.lib()
    pd = ctypes.POINTER(ctypes.c_double)
    buf = np.zeros(16)
    P = buf.ctypes.data_as(pd)
    s, a, b = ctypes.c_uint(), ctypes.c_ulonglong(), ctypes.c_ulonglong()
    return L.hypJacL2(routine, nrow, ncol, nplus, P if G else None,
                      nrow if ldhG is None else ldhG, P if V else None,
                      ncol if ldhV is None else ldhV, P if D else None,
                      ctypes.byref(s) if sw else None, ctypes.byref(a) if gs else None,
                      ctypes.byref(b) if gb else None, None)


@pytest.mark.parametrize("kw,code", [
    (dict(routine=16, nrow=256, ncol=256, nplus=256), -1),
    (dict(routine=9, nrow=0, ncol=256, nplus=256), -2),
    (dict(routine=9, nrow=96, ncol=64, nplus=64), -2),
    (dict(routine=9, nrow=256, ncol=0, nplus=0), -3),
    (dict(routine=9, nrow=256, ncol=512, nplus=0), -3),
    (dict(routine=9, nrow=256, ncol=48, nplus=0), -3),
    (dict(routine=9, nrow=256, ncol=64, nplus=65), -4),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, G=False), -5),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, ldhG=255), -6),
    (dict(routine=11, nrow=256, ncol=64, nplus=64, V=False), -7),
    (dict(routine=11, nrow=256, ncol=64, nplus=64, ldhV=63), -8),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, D=False), -9),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, sw=False), -10),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, gs=False), -11),
    (dict(routine=9, nrow=256, ncol=64, nplus=64, gb=False), -12),
])
def test_hypJacL2_argument_codes(kw, code):
    assert call(**kw) == code


def test_V_unchecked_without_full_svd():
    # the reference checks hV/ldhV only when routine & 2 (1/HypJacL2.cu:62-67)
    assert call(routine=9, nrow=256, ncol=64, nplus=64, V=False, ldhV=0, D=False) == -9


def test_gram_nchunk_matches_oracle():
    import oracle as o
    for nrow in (64, 128, 256, 512, 1024, 4096, 8192, 16384, 192, 320):
        for np_ in (1, 2, 8, 64, 128, 256, 512):
            assert g.gram_nchunk(nrow, np_) == o.gram_nchunk(nrow, np_)
