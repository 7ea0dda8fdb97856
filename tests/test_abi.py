"""The C-ABI library: it loads, exports every function that include/*.h
declares, and its host-side binned primitives (include/binned.h, which run on
the host as in the reference) agree bit for bit with the oracle. No GPU work
is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
        for m in re.finditer(r"\b((?:reproBLAS|binnedBLAS|binned|reproblas_hip)_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_d_path():
    names = set(declared_functions())
    for must in ("reproBLAS_rdsum", "reproBLAS_rddot", "reproBLAS_rdnrm2", "reproBLAS_rdasum", "reproBLAS_rdgemv",
                 "reproBLAS_rdgemm", "reproBLAS_dsum", "reproBLAS_ddot", "reproBLAS_dnrm2", "reproBLAS_dgemv",
                 "reproBLAS_dgemm", "binnedBLAS_dbdsum", "binnedBLAS_dbddot", "binnedBLAS_dbdssq",
                 "binnedBLAS_dbdgemv", "binnedBLAS_dbdgemm", "binned_dbdbadd", "binned_ddbconv", "binned_dbdconv"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import reproblas_amd

    L = reproblas_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def _L():
    import reproblas_amd

    L = reproblas_amd.lib()
    L.binned_dbdadd.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    L.binned_dbdbaddsq.restype = ctypes.c_double
    L.binned_dbdbaddsq.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    L.binned_dscale.restype = ctypes.c_double
    L.binned_dscale.argtypes = [ctypes.c_double]
    L.binned_dindex.restype = ctypes.c_int
    L.binned_dindex.argtypes = [ctypes.c_double]
    L.binned_dmbins.restype = ctypes.POINTER(ctypes.c_double)
    L.binned_dmbins.argtypes = [ctypes.c_int]
    L.binned_dbnegate.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.binned_dbrenorm.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.binned_dbbound.restype = ctypes.c_double
    L.binned_dbbound.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return L


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("fold", [2, 3, 4, 6, 8])
def test_host_binned_primitives_match_oracle(fold):
    """binned_dbdadd / dbdbadd / ddbconv / dbdbaddsq share binned_core.hpp with
    the kernels; here the host build of that code meets the oracle."""
    L, O = _L(), oracle()
    g = dg.rng(fold)
    for trial in range(400):
        a = g.normal(size=int(g.integers(1, 12))) * np.exp2(float(g.integers(-1070, 1015)))
        b = g.normal(size=int(g.integers(1, 12))) * np.exp2(float(g.integers(-1070, 1015)))
        Xo, Yo, Xg, Yg = (np.zeros(2 * fold) for _ in range(4))
        for v in a:
            O.dbdadd(fold, v, Xo)
            L.binned_dbdadd(fold, v, vp(Xg))
        for v in b:
            O.dbdadd(fold, v, Yo)
            L.binned_dbdadd(fold, v, vp(Yg))
        assert bits_equal(Xo, Xg) and bits_equal(Yo, Yg), trial
        sx, sy = O.dscale(abs(a[0])), O.dscale(abs(b[0]))
        Yo2, Yg2 = Yo.copy(), Yg.copy()
        assert O.dbdbaddsq(fold, sx, Xo, sy, Yo2) == L.binned_dbdbaddsq(fold, sx, vp(Xg), sy, vp(Yg2))
        assert bits_equal(Yo2, Yg2)
        O.dbdbadd(fold, Xo, Yo)
        L.binned_dbdbadd(fold, vp(Xg), vp(Yg))
        assert bits_equal(Yo, Yg)
        assert bits_equal([O.ddbconv(fold, Yo)], [L.binned_ddbconv(fold, vp(Yg))])


def test_host_scalar_helpers_match_oracle():
    L, O = _L(), oracle()
    for v in list(np.ldexp(1.0, np.arange(-1074, 1024, 1))) + [0.0, np.inf, np.nan, 3.0, -7.5]:
        assert L.binned_dindex(v) == O.dindex(v)
        assert bits_equal([L.binned_dscale(v)], [O.dscale(v)])
    ob = O.lib.orc_dmbins
    ob.restype = ctypes.POINTER(ctypes.c_double)
    ob.argtypes = [ctypes.c_int]
    for i in range(51 + 52):
        assert L.binned_dmbins(0)[i] == ob(0)[i]
    for fold in (2, 3, 5):
        assert L.binned_dbbound(fold, 1234, 2.5, 99.0) == O.dbbound(fold, 1234, 2.5, 99.0)


def test_negate_is_exact_inverse():
    L = _L()
    for fold in (2, 3, 4):
        X = np.zeros(2 * fold)
        for v in dg.wide(50, 3):
            L.binned_dbdadd(fold, v, vp(X))
        Y = X.copy()
        L.binned_dbnegate(fold, vp(Y))
        assert L.binned_ddbconv(fold, vp(Y)) == -L.binned_ddbconv(fold, vp(X))
        L.binned_dbdbadd(fold, vp(X), vp(Y))
        assert L.binned_ddbconv(fold, vp(Y)) == 0.0
