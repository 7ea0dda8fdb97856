"""Pins the complex oracle (oracle/oracle_z.c via oracle_cplx.inc; double
complex "z" and float complex "c") before it is trusted as the reference for
the complex kernels.

1. Bit-equality with the compiled reference (oracle/_ref: the reference's
   own z- and c-path C sources, AVX and scalar builds): sum, dotu, dotc,
   asum, nrm2 over random, ill-conditioned, extreme and non-finite inputs, N
   straddling the block edges, strides, folds; accumulate chains; gemv (both
   orders, N/T/C, complex alpha/beta including 0 and 1) and gemm (all 18
   layouts).
2. Properties that tie the complex path to the already pinned real oracle:
   with zero imaginary parts the real components equal the real path's
   results bit for bit; dzasum/dznrm2 equal the real asum/nrm2 of the 2N
   interleaved reals; conj(x).x has an exactly zero imaginary part;
   permutation invariance; the conjugation identity dotc(x, y) = dotu(conj x, y).
"""
import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle, oracle_s, oracle_z, reference_z

PRECS = ["z", "c"]


def bits_eq(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    assert a.dtype == b.dtype
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(u) == b.view(u))))


def _refs(prec):
    out = [r for r in (reference_z(prec, "avx"), reference_z(prec, "sisd")) if r is not None]
    if not out:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return out


def cvec(prec, kind, n, seed):
    """n complex values as 2n interleaved reals of the precision's type."""
    dt = np.float64 if prec == "z" else np.float32
    if kind == "wide":
        return dg.wide(2 * n, seed, -20, 20).astype(dt)
    if kind == "uniform":
        return dg.uniform(2 * n, seed).astype(dt)
    if kind == "ill":
        return dg.ill_conditioned(2 * n, seed).astype(dt)
    if kind == "real_only":
        v = dg.wide(2 * n, seed, -20, 20).astype(dt)
        v[1::2] = 0
        return v
    if prec == "z":
        return dg.special(kind, 2 * n, seed)
    return dg.special_s(kind, 2 * n, seed)


KINDS = ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "nan", "pinf", "ninf", "both_inf"]


def _state_eq(a, b):
    """Binned states: specified only where the component's top primary is finite."""
    fold = len(a) // 4
    ok = True
    for c in (0, 1):
        if np.isfinite(a[c]) and np.isfinite(b[c]):
            ok &= bits_eq(a[c:2 * fold:2], b[c:2 * fold:2]) and bits_eq(a[2 * fold + c::2], b[2 * fold + c::2])
        else:
            ok &= bits_eq(a[c], b[c]) or (np.isnan(a[c]) and np.isnan(b[c]))
    return ok


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("fold", [2, 3, 4, 7])
def test_blas1_matches_compiled_reference(prec, kind, fold):
    O = oracle_z(prec)
    for R in _refs(prec):
        for n in (1, 2, 255, 256, 257, 511, 512, 513, 1025, 2049, 3000):
            x, y = cvec(prec, kind, n, 10 + n), cvec(prec, "wide", n, 20 + n)
            assert bits_eq(O.rsum(fold, x), R.rsum(fold, x)), (kind, fold, n, "sum")
            for k in ("u", "c"):
                assert bits_eq(O.rdot(k, fold, x, y), R.rdot(k, fold, x, y)), (kind, fold, n, "dot" + k)
            assert bits_eq(O.rasum(fold, x), R.rasum(fold, x)), (kind, fold, n, "asum")
            assert bits_eq(O.rnrm2(fold, x), R.rnrm2(fold, x)), (kind, fold, n, "nrm2")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("fold", [2, 3, 9])
def test_states_and_accumulate_chains(prec, fold):
    O = oracle_z(prec)
    dt = O.dtype
    for R in _refs(prec):
        Ya, Yb = np.zeros(4 * fold, dt), np.zeros(4 * fold, dt)
        Da, Db = np.zeros(4 * fold, dt), np.zeros(4 * fold, dt)
        Ua, Ub = np.zeros(4 * fold, dt), np.zeros(4 * fold, dt)
        Aa, Ab = np.zeros(2 * fold, dt), np.zeros(2 * fold, dt)
        Qa, Qb = np.zeros(2 * fold, dt), np.zeros(2 * fold, dt)
        sa = sb = 0.0
        chunks = [cvec(prec, "wide", 900, 1), cvec(prec, "wide", 5000, 2) * dt(1e10),
                  cvec(prec, "uniform", 3, 3) * dt(1e-10), cvec(prec, "big", 100, 4), cvec(prec, "real_only", 700, 5)]
        for k, x in enumerate(chunks):
            y = cvec(prec, "wide", len(x) // 2, 50 + k)
            O.bsum(fold, x, Ya)
            R.bsum(fold, x, Yb)
            O.bdot("c", fold, x, y, Da)
            R.bdot("c", fold, x, y, Db)
            O.bdot("u", fold, x, y, Ua)
            R.bdot("u", fold, x, y, Ub)
            O.basum(fold, x, Aa)
            R.basum(fold, x, Ab)
            sa, sb = O.bssq(fold, x, sa, Qa), R.bssq(fold, x, sb, Qb)
            assert _state_eq(Ya, Yb) and _state_eq(Da, Db) and _state_eq(Ua, Ub), (prec, fold, k)
            assert bits_eq(Aa, Ab) and bits_eq(Qa, Qb) and bits_eq(np.array(sa, dt), np.array(sb, dt)), (fold, k)
        assert bits_eq(O.zbconv(fold, Ya), R.zbconv(fold, Yb))
        # binned add and scalar conversion
        Xa = O.bconv(fold, complex(3.5e7, -2.25e-3))
        Xb = R.bconv(fold, complex(3.5e7, -2.25e-3))
        assert bits_eq(Xa, Xb)
        O.fn["bbadd"](fold, O.p(Xa), O.p(Ya))
        R.fn["bbadd"](fold, R.p(Xb), R.p(Yb))
        assert _state_eq(Ya, Yb)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("inc", [2, -1, -3])
def test_strided_matches_reference(prec, inc):
    O = oracle_z(prec)
    n = 1501
    base = cvec(prec, "wide", n * abs(inc), 5)
    x, y = O.p(base), O.p(base)
    import ctypes

    # negative increments: the pointer is the lowest-addressed element (the
    # reference indexes X[2*i*incX] from it, walking backwards)
    off = 0 if inc > 0 else (n - 1) * abs(inc) * 2 * base.itemsize
    ptr = ctypes.c_void_p(base.ctypes.data + off)
    for R in _refs(prec):
        for k in ("rasum", "rnrm2"):
            assert bits_eq(np.array(O.fn[k](3, n, ptr, inc), O.dtype), np.array(R.fn[k](3, n, ptr, inc), O.dtype))
        for k in ("rsum",):
            a, b = np.zeros(2, O.dtype), np.zeros(2, O.dtype)
            O.fn[k](3, n, ptr, inc, O.p(a))
            R.fn[k](3, n, ptr, inc, R.p(b))
            assert bits_eq(a, b)
        for k in ("rdotu", "rdotc"):
            a, b = np.zeros(2, O.dtype), np.zeros(2, O.dtype)
            O.fn[k](3, n, ptr, inc, ptr, inc, O.p(a))
            R.fn[k](3, n, ptr, inc, ptr, inc, R.p(b))
            assert bits_eq(a, b), k
    del x, y


@pytest.mark.parametrize("prec", PRECS)
def test_amax_matches_reference(prec):
    O = oracle_z(prec)
    for R in _refs(prec):
        for kind in ("wide", "nan", "pinf"):
            x, y = cvec(prec, kind, 777, 3), cvec(prec, "wide", 777, 4)
            a, b = np.zeros(2, O.dtype), np.zeros(2, O.dtype)
            O.fn["amax"](777, O.p(x), 1, O.p(a))
            R.fn["amax"](777, R.p(x), 1, R.p(b))
            assert bits_eq(a, b)
            O.fn["amaxm"](777, O.p(x), 1, O.p(y), 1, O.p(a))
            R.fn["amaxm"](777, R.p(x), 1, R.p(y), 1, R.p(b))
            assert bits_eq(a, b)


ALPHAS = [complex(1, 0), complex(2, -0.5), complex(0, 1), complex(0, 0)]
BETAS = [complex(0, 0), complex(1, 0), complex(0.5, 0.25)]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_gemv_matches_reference(prec, order, trans):
    O = oracle_z(prec)
    # the reference's column-major conjugate-transpose branch dots N - j rows
    # of each M-row column (zbzgemv.ccog:218): defined for M >= N (the first N
    # rows), past the column for M < N, so that branch takes M >= N only
    shapes = [(57, 57), (131, 97), (300, 40)] if (order == "C" and trans == "C") else [(131, 97), (40, 300)]
    for M, N in shapes:
        lda = (M if order == "C" else N) + 2
        A = cvec(prec, "wide", lda * max(M, N), 31)
        x = cvec(prec, "uniform", max(M, N), 32)
        y0 = cvec(prec, "uniform", max(M, N), 33)
        for R in _refs(prec):
            for fold in (2, 3):
                for alpha in ALPHAS:
                    for beta in BETAS:
                        ya, yb = y0.copy(), y0.copy()
                        al, be = O.scalar(alpha), O.scalar(beta)  # kept alive across the calls
                        for L, y in ((O, ya), (R, yb)):
                            L.fn["rgemv"](fold, order.encode(), trans.encode(), M, N, L.p(al), L.p(A), lda, L.p(x), 1,
                                          L.p(be), L.p(y), 1)
                        assert bits_eq(ya, yb), (M, N, fold, alpha, beta)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
def test_gemm_matches_reference(prec, order, ta, tb):
    O = oracle_z(prec)
    M, N, K = 23, 17, 300
    rm = order == "R"
    lda = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
    ldb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
    ldc = N if rm else M
    A = cvec(prec, "wide", K * max(M, N) + 8, 41)
    B = cvec(prec, "wide", K * max(M, N) + 8, 42)
    C0 = cvec(prec, "uniform", M * N + 8, 43)
    for R in _refs(prec):
        for alpha, beta in ((complex(1, 0), complex(0, 0)), (complex(1.5, -2), complex(0.5, 0.5)),
                            (complex(1, 0), complex(1, 0)), (complex(0, 0), complex(2, 0))):
            ca, cb = C0.copy(), C0.copy()
            al, be = O.scalar(alpha), O.scalar(beta)  # kept alive across the calls
            for L, Cm in ((O, ca), (R, cb)):
                L.fn["rgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, L.p(al), L.p(A), lda, L.p(B), ldb,
                              L.p(be), L.p(Cm), ldc)
            assert bits_eq(ca, cb), (alpha, beta)


# ------------------------------------------------------------ properties --
@pytest.mark.parametrize("prec", PRECS)
def test_reduces_to_real_path(prec):
    """Zero imaginary parts: the real component equals the real path's result
    bit for bit; asum/nrm2 equal the real path over the 2N interleaved reals."""
    Oz = oracle_z(prec)
    Or = oracle() if prec == "z" else oracle_s()
    for fold in (2, 3, 4):
        for n in (1, 1000, 5000):
            x = cvec(prec, "real_only", n, 70 + n)
            y = cvec(prec, "real_only", n, 80 + n)
            xr, yr = np.ascontiguousarray(x[0::2]), np.ascontiguousarray(y[0::2])
            s = Oz.rsum(fold, x)
            rs = Or.rdsum(fold, xr) if prec == "z" else Or.rssum(fold, xr)
            assert bits_eq(s[0], np.array(rs, Oz.dtype)) and s[1] == 0
            d = Oz.rdot("u", fold, x, y)
            rd = Or.rddot(fold, xr, yr) if prec == "z" else Or.rsdot(fold, xr, yr)
            assert bits_eq(d[0], np.array(rd, Oz.dtype)), (fold, n)
            w = cvec(prec, "wide", n, 90 + n)
            ra = Or.rdasum(fold, w) if prec == "z" else Or.rsasum(fold, w)
            rn = Or.rdnrm2(fold, w) if prec == "z" else Or.rsnrm2(fold, w)
            assert bits_eq(np.array(Oz.rasum(fold, w), Oz.dtype), np.array(ra, Oz.dtype)), (fold, n)
            assert bits_eq(np.array(Oz.rnrm2(fold, w), Oz.dtype), np.array(rn, Oz.dtype)), (fold, n)


@pytest.mark.parametrize("prec", PRECS)
def test_conjugation_identities(prec):
    O = oracle_z(prec)
    for fold in (2, 3):
        x, y = cvec(prec, "wide", 3000, 1), cvec(prec, "wide", 3000, 2)
        assert O.rdot("c", fold, x, x)[1] == 0  # conj(x).x is real: its imaginary terms cancel in pairs
        xc = x.copy()
        xc[1::2] = -xc[1::2]
        assert bits_eq(O.rdot("c", fold, x, y), O.rdot("u", fold, xc, y))
        perm = dg.rng(5).permutation(3000)
        xp = x.reshape(-1, 2)[perm].ravel().copy()
        yp = y.reshape(-1, 2)[perm].ravel().copy()
        assert bits_eq(O.rsum(fold, xp), O.rsum(fold, x))
        assert bits_eq(O.rdot("u", fold, xp, yp), O.rdot("u", fold, x, y))


# ---------------------------------------------------------------- golden --
@pytest.mark.parametrize("prec", PRECS)
def test_oracle_matches_golden_fixtures(prec):
    """tests/golden/sc.npz holds the compiled reference's complex outputs
    (tests/golden/make_golden_sc.py); the oracle reproduces them without the
    reference library."""
    import os

    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sc.npz"))
    O = oracle_z(prec)
    keys = [k for k in G.files if k.startswith(prec + "_") and k.endswith("_res")]
    assert keys
    for k in keys:
        parts = k.split("_")
        kind, n, fold = "_".join(parts[1:-3]), parts[-3], int(parts[-2])
        x, y = G[f"{prec}_{kind}_{n}_x"], G[f"{prec}_{kind}_{n}_y"]
        got = np.concatenate([O.rsum(fold, x), O.rdot("u", fold, x, y), O.rdot("c", fold, x, y),
                              np.array([O.rasum(fold, x), O.rnrm2(fold, x)], x.dtype)])
        assert bits_eq(got, G[k]), k
    A, B, C = G[f"{prec}_mat_A"], G[f"{prec}_mat_B"], G[f"{prec}_mat_C"]
    dt = A.dtype
    M, N, K = 23, 17, 200
    al, be = np.array([1.5, -0.5], dt), np.array([0.5, 0.25], dt)
    for order in "CR":
        for ta in "NTC":
            if not (order == "C" and ta == "C"):
                y = C.copy()
                lda = (M if order == "C" else N) + 1
                O.fn["rgemv"](3, order.encode(), ta.encode(), M, N, O.p(al), O.p(A), lda, O.p(B), 1, O.p(be), O.p(y), 1)
                assert bits_eq(y, G[f"{prec}_gemv_{order}{ta}"]), (order, ta)
            for tb in "NTC":
                rm = order == "R"
                la = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
                lb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
                c = C.copy()
                O.fn["rgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, O.p(al), O.p(A), la, O.p(B), lb,
                              O.p(be), O.p(c), N if rm else M)
                assert bits_eq(c, G[f"{prec}_gemm_{order}{ta}{tb}"]), (order, ta, tb)
