"""ctypes handles on the CPU oracle libraries (test infrastructure only).

* ``oracle()``  -- oracle/liboracle.so, the scalar restatement (orc_* symbols);
  ``oracle_s()`` the same library's single-precision calls.
* ``reference(kind)`` -- oracle/_ref/libreproblas_ref{,_sisd}.so, the
  reference's own C sources compiled by oracle/Makefile; ``None`` when they
  were not built (the reference tree is only present where the fixtures are
  made); ``reference_s(kind)`` its single-precision calls.

Both are wrapped in :class:`CpuBlas`, which exposes the reproBLAS-style calls
on numpy arrays so a test can run the same call on either library.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = {
    "avx": os.path.join(ROOT, "oracle", "_ref", "libreproblas_ref.so"),
    "sisd": os.path.join(ROOT, "oracle", "_ref", "libreproblas_ref_sisd.so"),
}

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_char = ctypes.c_char
_vp = ctypes.c_void_p


def _ptr(a):
    if not a.flags["C_CONTIGUOUS"] or a.dtype != np.float64:
        raise ValueError("oracle calls need contiguous float64 arrays")
    return ctypes.c_void_p(a.ctypes.data)


class CpuBlas:
    """reproBLAS calls on numpy float64 arrays through one CPU library."""

    def __init__(self, path, style):
        self.path = path
        self.style = style  # "orc" or "ref"
        self.lib = ctypes.CDLL(path)
        n = self._name
        sig = {
            "rdsum": (_c_double, [_c_int, _c_int, _vp, _c_int]),
            "rdasum": (_c_double, [_c_int, _c_int, _vp, _c_int]),
            "rdnrm2": (_c_double, [_c_int, _c_int, _vp, _c_int]),
            "rddot": (_c_double, [_c_int, _c_int, _vp, _c_int, _vp, _c_int]),
            "rdgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_double, _vp, _c_int,
                              _vp, _c_int, _c_double, _vp, _c_int]),
            "rdgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_double,
                              _vp, _c_int, _vp, _c_int, _c_double, _vp, _c_int]),
            "dbdsum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
            "dbdasum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
            "dbddot": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp]),
            "dbdssq": (_c_double, [_c_int, _c_int, _vp, _c_int, _c_double, _vp]),
            "dbdgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_double, _vp, _c_int,
                               _vp, _c_int, _vp, _c_int]),
            "dbdgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_double,
                               _vp, _c_int, _vp, _c_int, _vp, _c_int]),
            "dbdbadd": (None, [_c_int, _vp, _vp]),
            "dbdbaddsq": (_c_double, [_c_int, _c_double, _vp, _c_double, _vp]),
            "dbdconv": (None, [_c_int, _c_double, _vp]),
            "ddbconv": (_c_double, [_c_int, _vp]),
            "dbdadd": (None, [_c_int, _c_double, _vp]),
            "dindex": (_c_int, [_c_double]),
            "dmindex": (_c_int, [_vp]),
            "dscale": (_c_double, [_c_double]),
            "dbbound": (_c_double, [_c_int, _c_int, _c_double, _c_double]),
            "damax": (_c_double, [_c_int, _vp, _c_int]),
            "damaxm": (_c_double, [_c_int, _vp, _c_int, _vp, _c_int]),
        }
        self.fn = {}
        for k, (res, args) in sig.items():
            f = getattr(self.lib, n(k))
            f.restype = res
            f.argtypes = args
            self.fn[k] = f

    def _name(self, k):
        if self.style == "orc":
            return "orc_" + k
        if k in ("rdsum", "rdasum", "rdnrm2", "rddot", "rdgemv", "rdgemm"):
            return "reproBLAS_" + k
        if k in ("dbdsum", "dbdasum", "dbddot", "dbdssq", "dbdgemv", "dbdgemm", "damax", "damaxm"):
            return "binnedBLAS_" + k
        return "binned_" + k

    # --- BLAS1 -----------------------------------------------------------
    def rdsum(self, fold, x, inc=1, n=None):
        n = len(x) // max(abs(inc), 1) if n is None else n
        return self.fn["rdsum"](fold, n, _ptr(x), inc)

    def rdasum(self, fold, x, inc=1, n=None):
        n = len(x) // max(abs(inc), 1) if n is None else n
        return self.fn["rdasum"](fold, n, _ptr(x), inc)

    def rdnrm2(self, fold, x, inc=1, n=None):
        n = len(x) // max(abs(inc), 1) if n is None else n
        return self.fn["rdnrm2"](fold, n, _ptr(x), inc)

    def rddot(self, fold, x, y, incx=1, incy=1, n=None):
        n = min(len(x) // max(abs(incx), 1), len(y) // max(abs(incy), 1)) if n is None else n
        return self.fn["rddot"](fold, n, _ptr(x), incx, _ptr(y), incy)

    # --- binned-state BLAS1 (accumulate into Y, a float64 array of 2*fold) --
    def dbdsum(self, fold, x, Y, inc=1):
        self.fn["dbdsum"](fold, len(x) // abs(inc), _ptr(x), inc, _ptr(Y))

    def dbdasum(self, fold, x, Y, inc=1):
        self.fn["dbdasum"](fold, len(x) // abs(inc), _ptr(x), inc, _ptr(Y))

    def dbddot(self, fold, x, y, Z):
        self.fn["dbddot"](fold, len(x), _ptr(x), 1, _ptr(y), 1, _ptr(Z))

    def dbdssq(self, fold, x, scale, Y):
        return self.fn["dbdssq"](fold, len(x), _ptr(x), 1, scale, _ptr(Y))

    # --- primitives ---------------------------------------------------------
    def dbdbadd(self, fold, X, Y):
        self.fn["dbdbadd"](fold, _ptr(X), _ptr(Y))

    def dbdbaddsq(self, fold, sx, X, sy, Y):
        return self.fn["dbdbaddsq"](fold, sx, _ptr(X), sy, _ptr(Y))

    def dbdconv(self, fold, x):
        Y = np.zeros(2 * fold)
        self.fn["dbdconv"](fold, x, _ptr(Y))
        return Y

    def ddbconv(self, fold, Y):
        return self.fn["ddbconv"](fold, _ptr(Y))

    def dbdadd(self, fold, x, Y):
        self.fn["dbdadd"](fold, x, _ptr(Y))

    def dindex(self, x):
        return self.fn["dindex"](x)

    def dmindex(self, Y):
        return self.fn["dmindex"](_ptr(Y))

    def dscale(self, x):
        return self.fn["dscale"](x)

    def dbbound(self, fold, n, x, s):
        return self.fn["dbbound"](fold, n, x, s)

    # --- level 2/3 ----------------------------------------------------------
    def rdgemv(self, fold, order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy):
        self.fn["rdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, _ptr(A), lda, _ptr(x), incx,
                          beta, _ptr(y), incy)

    def rdgemm(self, fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
        self.fn["rdgemm"](fold, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A), lda,
                          _ptr(B), ldb, beta, _ptr(C), ldc)

    def dbdgemm(self, fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc):
        self.fn["dbdgemm"](fold, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A), lda,
                           _ptr(B), ldb, _ptr(C), ldc)


_c_float = ctypes.c_float


def _fptr(a):
    if not a.flags["C_CONTIGUOUS"] or a.dtype != np.float32:
        raise ValueError("oracle calls need contiguous float32 arrays")
    return ctypes.c_void_p(a.ctypes.data)


class CpuBlasS:
    """The single-precision (float_binned) calls on numpy float32 arrays."""

    def __init__(self, path, style):
        self.path = path
        self.style = style
        self.lib = ctypes.CDLL(path)
        f, i, c, v = _c_float, _c_int, _c_char, _vp
        sig = {
            "rssum": (f, [i, i, v, i]), "rsasum": (f, [i, i, v, i]), "rsnrm2": (f, [i, i, v, i]),
            "rsdot": (f, [i, i, v, i, v, i]),
            "rsgemv": (None, [i, c, c, i, i, f, v, i, v, i, f, v, i]),
            "rsgemm": (None, [i, c, c, c, i, i, i, f, v, i, v, i, f, v, i]),
            "sbssum": (None, [i, i, v, i, v]), "sbsasum": (None, [i, i, v, i, v]),
            "sbsdot": (None, [i, i, v, i, v, i, v]), "sbsssq": (f, [i, i, v, i, f, v]),
            "sbsgemv": (None, [i, c, c, i, i, f, v, i, v, i, v, i]),
            "sbsgemm": (None, [i, c, c, c, i, i, i, f, v, i, v, i, v, i]),
            "sbsbadd": (None, [i, v, v]), "sbsbaddsq": (f, [i, f, v, f, v]), "sbsconv": (None, [i, f, v]),
            "ssbconv": (f, [i, v]), "sbsadd": (None, [i, f, v]), "sindex": (i, [f]), "smindex": (i, [v]),
            "sscale": (f, [f]), "sbbound": (f, [i, i, f, f]), "samax": (f, [i, v, i]),
            "samaxm": (f, [i, v, i, v, i]),
        }
        self.fn = {}
        for k, (res, args) in sig.items():
            fn = getattr(self.lib, self._name(k))
            fn.restype = res
            fn.argtypes = args
            self.fn[k] = fn

    def _name(self, k):
        if self.style == "orc":
            return "orc_" + k
        if k.startswith("rs"):
            return "reproBLAS_" + k
        if k in ("sbssum", "sbsasum", "sbsdot", "sbsssq", "sbsgemv", "sbsgemm", "samax", "samaxm"):
            return "binnedBLAS_" + k
        return "binned_" + k

    def rssum(self, fold, x, inc=1):
        return self.fn["rssum"](fold, len(x) // abs(inc), _fptr(x), inc)

    def rsasum(self, fold, x, inc=1):
        return self.fn["rsasum"](fold, len(x) // abs(inc), _fptr(x), inc)

    def rsnrm2(self, fold, x, inc=1):
        return self.fn["rsnrm2"](fold, len(x) // abs(inc), _fptr(x), inc)

    def rsdot(self, fold, x, y):
        return self.fn["rsdot"](fold, min(len(x), len(y)), _fptr(x), 1, _fptr(y), 1)

    def sbssum(self, fold, x, Y):
        self.fn["sbssum"](fold, len(x), _fptr(x), 1, _fptr(Y))

    def sbsasum(self, fold, x, Y):
        self.fn["sbsasum"](fold, len(x), _fptr(x), 1, _fptr(Y))

    def sbsdot(self, fold, x, y, Z):
        self.fn["sbsdot"](fold, len(x), _fptr(x), 1, _fptr(y), 1, _fptr(Z))

    def sbsssq(self, fold, x, scale, Y):
        return self.fn["sbsssq"](fold, len(x), _fptr(x), 1, scale, _fptr(Y))

    def sbsbadd(self, fold, X, Y):
        self.fn["sbsbadd"](fold, _fptr(X), _fptr(Y))

    def sbsadd(self, fold, x, Y):
        self.fn["sbsadd"](fold, x, _fptr(Y))

    def ssbconv(self, fold, Y):
        return self.fn["ssbconv"](fold, _fptr(Y))


_cache = {}


def oracle():
    if "orc" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        _cache["orc"] = CpuBlas(ORACLE_SO, "orc")
    return _cache["orc"]


def reference(kind="avx"):
    key = "ref_" + kind
    if key not in _cache:
        p = REF_SO[kind]
        _cache[key] = CpuBlas(p, "ref") if os.path.exists(p) else None
    return _cache[key]


class CpuBlasZ:
    """The complex calls of one precision ("z": complex128 data as float64
    pairs, "c": complex64 as float32 pairs) on interleaved numpy arrays."""

    # key -> (oracle name, reference name) with {z} (z/c), {d} (d/s), {rd} (rdz/rsc)
    NAMES = {
        "rsum": ("orc_r{z}sum", "reproBLAS_r{z}sum_sub"),
        "rdotu": ("orc_r{z}dotu", "reproBLAS_r{z}dotu_sub"),
        "rdotc": ("orc_r{z}dotc", "reproBLAS_r{z}dotc_sub"),
        "rasum": ("orc_{rd}asum", "reproBLAS_{rd}asum"),
        "rnrm2": ("orc_{rd}nrm2", "reproBLAS_{rd}nrm2"),
        "rgemv": ("orc_r{z}gemv", "reproBLAS_r{z}gemv"),
        "rgemm": ("orc_r{z}gemm", "reproBLAS_r{z}gemm"),
        "bsum": ("orc_{z}b{z}sum", "binnedBLAS_{z}b{z}sum"),
        "bdotu": ("orc_{z}b{z}dotu", "binnedBLAS_{z}b{z}dotu"),
        "bdotc": ("orc_{z}b{z}dotc", "binnedBLAS_{z}b{z}dotc"),
        "basum": ("orc_{d}b{z}asum", "binnedBLAS_{d}b{z}asum"),
        "bssq": ("orc_{d}b{z}ssq", "binnedBLAS_{d}b{z}ssq"),
        "bgemv": ("orc_{z}b{z}gemv", "binnedBLAS_{z}b{z}gemv"),
        "bgemm": ("orc_{z}b{z}gemm", "binnedBLAS_{z}b{z}gemm"),
        "bbadd": ("orc_{z}b{z}badd", "binned_{z}b{z}badd"),
        "bconv": ("orc_{z}b{z}conv", "binned_{z}b{z}conv"),
        "zbconv": ("orc_{z}{z}bconv", "binned_{z}{z}bconv_sub"),
        "amax": ("orc_{z}amax", "binnedBLAS_{z}amax_sub"),
        "amaxm": ("orc_{z}amaxm", "binnedBLAS_{z}amaxm_sub"),
    }

    def __init__(self, path, style, prec):
        self.path, self.style, self.prec = path, style, prec
        self.lib = ctypes.CDLL(path)
        self.dtype = np.float64 if prec == "z" else np.float32
        R = _c_double if prec == "z" else _c_float
        i, c, v = _c_int, _c_char, _vp
        sig = {
            "rsum": (None, [i, i, v, i, v]), "rdotu": (None, [i, i, v, i, v, i, v]),
            "rdotc": (None, [i, i, v, i, v, i, v]), "rasum": (R, [i, i, v, i]), "rnrm2": (R, [i, i, v, i]),
            "rgemv": (None, [i, c, c, i, i, v, v, i, v, i, v, v, i]),
            "rgemm": (None, [i, c, c, c, i, i, i, v, v, i, v, i, v, v, i]),
            "bsum": (None, [i, i, v, i, v]), "bdotu": (None, [i, i, v, i, v, i, v]),
            "bdotc": (None, [i, i, v, i, v, i, v]), "basum": (None, [i, i, v, i, v]),
            "bssq": (R, [i, i, v, i, R, v]),
            "bgemv": (None, [i, c, c, i, i, v, v, i, v, i, v, i]),
            "bgemm": (None, [i, c, c, c, i, i, i, v, v, i, v, i, v, i]),
            "bbadd": (None, [i, v, v]), "bconv": (None, [i, v, v]), "zbconv": (None, [i, v, v]),
            "amax": (None, [i, v, i, v]), "amaxm": (None, [i, v, i, v, i, v]),
        }
        sub = {"z": prec, "d": "d" if prec == "z" else "s", "rd": "rdz" if prec == "z" else "rsc"}
        self.fn = {}
        for k, (res, args) in sig.items():
            name = self.NAMES[k][0 if style == "orc" else 1].format(**sub)
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args
            self.fn[k] = f

    def p(self, a):
        if not a.flags["C_CONTIGUOUS"] or a.dtype != self.dtype:
            raise ValueError("complex oracle calls need contiguous interleaved arrays of the precision's type")
        return ctypes.c_void_p(a.ctypes.data)

    def scalar(self, v):
        return np.array([np.real(v), np.imag(v)], dtype=self.dtype)

    # x, y: interleaved arrays of 2n reals
    def rsum(self, fold, x, inc=1):
        out = np.zeros(2, self.dtype)
        self.fn["rsum"](fold, len(x) // (2 * abs(inc)), self.p(x), inc, self.p(out))
        return out

    def rdot(self, kind, fold, x, y):
        out = np.zeros(2, self.dtype)
        self.fn["rdot" + kind](fold, min(len(x), len(y)) // 2, self.p(x), 1, self.p(y), 1, self.p(out))
        return out

    def rasum(self, fold, x, inc=1):
        return self.fn["rasum"](fold, len(x) // (2 * abs(inc)), self.p(x), inc)

    def rnrm2(self, fold, x, inc=1):
        return self.fn["rnrm2"](fold, len(x) // (2 * abs(inc)), self.p(x), inc)

    def bsum(self, fold, x, Y):
        self.fn["bsum"](fold, len(x) // 2, self.p(x), 1, self.p(Y))

    def bdot(self, kind, fold, x, y, Z):
        self.fn["bdot" + kind](fold, len(x) // 2, self.p(x), 1, self.p(y), 1, self.p(Z))

    def basum(self, fold, x, Y):
        self.fn["basum"](fold, len(x) // 2, self.p(x), 1, self.p(Y))

    def bssq(self, fold, x, scale, Y):
        return self.fn["bssq"](fold, len(x) // 2, self.p(x), 1, scale, self.p(Y))

    def zbconv(self, fold, Y):
        out = np.zeros(2, self.dtype)
        self.fn["zbconv"](fold, self.p(Y), self.p(out))
        return out

    def bconv(self, fold, v):
        Y = np.zeros(4 * fold, self.dtype)
        sv = self.scalar(v)
        self.fn["bconv"](fold, self.p(sv), self.p(Y))
        return Y


def oracle_z(prec="z"):
    key = "orc_" + prec
    if key not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        _cache[key] = CpuBlasZ(ORACLE_SO, "orc", prec)
    return _cache[key]


def reference_z(prec="z", kind="avx"):
    key = f"ref_{prec}_{kind}"
    if key not in _cache:
        p = REF_SO[kind]
        _cache[key] = CpuBlasZ(p, "ref", prec) if os.path.exists(p) else None
    return _cache[key]


def oracle_s():
    if "orc_s" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        _cache["orc_s"] = CpuBlasS(ORACLE_SO, "orc")
    return _cache["orc_s"]


def reference_s(kind="avx"):
    key = "ref_s_" + kind
    if key not in _cache:
        p = REF_SO[kind]
        _cache[key] = CpuBlasS(p, "ref") if os.path.exists(p) else None
    return _cache[key]
