"""Pins the oracle (oracle/oracle.c) before it is trusted as the reference for the GPU results.

1. Bit-equality with the compiled reference (oracle/_ref: the reference's own
   d-path C sources, AVX and scalar builds) over random, ill-conditioned,
   extreme and non-finite inputs, strides, folds 2..8, gemv and gemm.
2. Bit-equality with the committed golden fixtures (tests/golden/), which the
   reference library wrote -- this part runs wherever the fixtures are.
3. Properties that the mathematics fixes, independent of any implementation:
   the reference's bound (binned_dbbound, README.md:65-71) against exact
   rational sums, permutation and blocking invariance
   (tests/checks/verify_daugsum.c), closed-form fills
   (tests/checks/wrap_daugsum.h:232-461), the index/scale bracket properties
   (validate_internal_dindex.c, validate_internal_dscale.c), single-element
   round trips, gemm == one binned dot per output (wrap_rdgemm.h:16-70).
"""
import ctypes
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle, reference

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return np.array([a]).view(np.uint64)[0] == np.array([b]).view(np.uint64)[0]


def eq_arrays(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint64) == b.view(np.uint64))))


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- reference --
def _refs():
    out = [r for r in (reference("avx"), reference("sisd")) if r is not None]
    if not out:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return out


INPUTS = ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "constant", "nan", "pinf", "ninf",
          "both_inf", "inf_nan"]


def _vec(kind, n, seed):
    if kind == "uniform":
        return dg.uniform(n, seed)
    if kind == "wide":
        return dg.wide(n, seed)
    if kind == "ill":
        return dg.ill_conditioned(n, seed)
    if n < 2 and kind in ("both_inf", "inf_nan"):
        return dg.special("nan", n, seed)
    return dg.special(kind, n, seed)


@pytest.mark.parametrize("kind", INPUTS)
@pytest.mark.parametrize("fold", [2, 3, 4, 5, 8, 13, 52])
def test_blas1_matches_compiled_reference(kind, fold):
    O = oracle()
    for R in _refs():
        for n in (1, 2, 3, 1023, 1024, 1025, 2047, 2049, 6000):
            x, y = _vec(kind, n, 10 + n), dg.wide(n, 20 + n)
            for op in ("rdsum", "rdasum", "rdnrm2"):
                assert same(getattr(O, op)(fold, x), getattr(R, op)(fold, x)), (kind, fold, n, op)
            assert same(O.rddot(fold, x, y), R.rddot(fold, x, y)), (kind, fold, n)
            for st_op in ("dbdsum", "dbdasum"):
                a, b = np.zeros(2 * fold), np.zeros(2 * fold)
                getattr(O, st_op)(fold, x, a)
                getattr(R, st_op)(fold, x, b)
                if not np.isnan(a[0]):
                    assert eq_arrays(a, b), (kind, fold, n, st_op)
            a, b = np.zeros(2 * fold), np.zeros(2 * fold)
            sa, sb = O.dbdssq(fold, x, 0.0, a), R.dbdssq(fold, x, 0.0, b)
            assert same(sa, sb)
            if not np.isnan(a[0]):
                assert eq_arrays(a, b)


@pytest.mark.parametrize("inc", [2, -1, -3])
def test_strided_matches_reference(inc):
    O = oracle()
    n = 3001
    base = dg.wide(n * abs(inc), 5)
    ptr = ctypes.c_void_p(base.ctypes.data + (0 if inc > 0 else (n - 1) * abs(inc) * 8))
    for R in _refs():
        for op in ("rdsum", "rdasum", "rdnrm2"):
            assert same(O.fn[op](3, n, ptr, inc), R.fn[op](3, n, ptr, inc))
        assert same(O.fn["rddot"](3, n, ptr, inc, ptr, inc), R.fn["rddot"](3, n, ptr, inc, ptr, inc))


def test_accumulate_chain_matches_reference():
    """Accumulating several vectors into one binned value (and one scaled
    ssq) gives the reference's state after each step."""
    O = oracle()
    for R in _refs():
        for fold in (2, 3, 4):
            a, b = np.zeros(2 * fold), np.zeros(2 * fold)
            sa = sb = 0.0
            qa, qb = np.zeros(2 * fold), np.zeros(2 * fold)
            for k, x in enumerate([dg.wide(900, 1), dg.wide(5000, 2) * 1e20, dg.uniform(3, 3) * 1e-20,
                                   dg.special("big", 100, 4)]):
                O.dbdsum(fold, x, a)
                R.dbdsum(fold, x, b)
                assert eq_arrays(a, b), (fold, k)
                sa, sb = O.dbdssq(fold, x, sa, qa), R.dbdssq(fold, x, sb, qb)
                assert same(sa, sb) and eq_arrays(qa, qb), (fold, k)


def test_primitives_match_reference():
    O = oracle()
    g = dg.rng(77)
    for R in _refs():
        for v in list(np.ldexp(1.0, np.arange(-1074, 1024, 3))) + [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324]:
            assert O.dindex(v) == R.dindex(v)
            assert same(O.dscale(v), R.dscale(v))
        for fold in (2, 3, 4, 7):
            for _ in range(300):
                xs = g.normal(size=5) * np.exp2(g.integers(-1060, 1010))
                ys = g.normal(size=5) * np.exp2(g.integers(-1060, 1010))
                X1, X2, Y1, Y2 = (np.zeros(2 * fold) for _ in range(4))
                for v in xs:
                    O.dbdadd(fold, v, X1)
                    R.dbdadd(fold, v, X2)
                for v in ys:
                    O.dbdadd(fold, v, Y1)
                    R.dbdadd(fold, v, Y2)
                assert eq_arrays(X1, X2) and eq_arrays(Y1, Y2)
                assert same(O.ddbconv(fold, X1), R.ddbconv(fold, X2))
                O.dbdbadd(fold, X1, Y1)
                R.dbdbadd(fold, X2, Y2)
                assert eq_arrays(Y1, Y2)
                s1, s2 = O.dscale(abs(xs[0])), O.dscale(abs(ys[0]))
                A1, A2 = Y1.copy(), Y2.copy()
                assert same(O.dbdbaddsq(fold, s1, X1, s2, A1), R.dbdbaddsq(fold, s1, X2, s2, A2))
                assert eq_arrays(A1, A2)
            assert same(O.dbbound(fold, 1000, 3.5, 17.0), R.dbbound(fold, 1000, 3.5, 17.0))


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_gemv_matches_reference(order, trans):
    O = oracle()
    M, N = 131, 97
    lda = (M if order == "C" else N) + 2
    A = dg.wide(lda * max(M, N), 31)
    x = dg.normal(max(M, N), 32)
    y0 = dg.normal(max(M, N), 33)
    for R in _refs():
        for fold in (2, 3):
            for alpha, beta in ((1.0, 0.0), (2.0, 0.5), (0.0, 1.0), (0.5, 1.0)):
                ya, yb = y0.copy(), y0.copy()
                for L, y in ((O, ya), (R, yb)):
                    L.fn["rdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                   vp(y), 1)
                assert eq_arrays(ya, yb), (fold, alpha, beta)


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_matches_reference(order, ta, tb):
    O = oracle()
    M, N, K = 37, 23, 1100
    A = dg.wide(K * max(M, N) + 64, 41)
    B = dg.wide(K * max(M, N) + 64, 42)
    C0 = dg.normal(max(M, N) ** 2 + 64, 43)
    rm = order == "R"
    lda = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
    ldb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
    ldc = N if rm else M
    for R in _refs():
        for alpha, beta in ((1.0, 0.0), (3.0, 0.5), (1.0, 1.0), (0.0, 2.0)):
            ca, cb = C0.copy(), C0.copy()
            for L, C in ((O, ca), (R, cb)):
                L.fn["rdgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb,
                               beta, vp(C), ldc)
            assert eq_arrays(ca, cb), (alpha, beta)


# ------------------------------------------------------------------ golden --
def _golden(name):
    p = os.path.join(GOLDEN, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} missing")
    return np.load(p, allow_pickle=False)


def test_oracle_matches_golden_blas1():
    O = oracle()
    G = _golden("blas1.npz")
    for key in G["cases"]:
        key = str(key)
        x = G["x_" + key]
        y = G["y_N" + key.rsplit("_N", 1)[1]]
        for fold in (2, 3, 4):
            for op in ("rdsum", "rdasum", "rdnrm2"):
                assert same(getattr(O, op)(fold, x), G[f"r_{op}_f{fold}_{key}"][0]), (key, op, fold)
            assert same(O.rddot(fold, x, y), G[f"r_rddot_f{fold}_{key}"][0]), (key, fold)
            st = np.zeros(2 * fold)
            O.dbdsum(fold, x, st)
            exp = G[f"b_dbdsum_f{fold}_{key}"]
            if not np.isnan(exp[0]):
                assert eq_arrays(st, exp), (key, fold)


def test_oracle_matches_golden_level23():
    O = oracle()
    G = _golden("level23.npz")
    for key in G["cases"]:
        key = str(key)
        if key.startswith("gemv"):
            _, ot, a, b, f = key.split("_")
            order, trans = ot[0], ot[1]
            alpha, beta, fold = float(a[1:]), float(b[1:]), int(f[1:])
            M, N = 257, 129
            lda = (M if order == "C" else N) + 1
            A = G["A_gemv_C"][: lda * (N if order == "C" else M)]
            x, y = G[f"x_gemv_{ot}"], G[f"y0_gemv_{ot}"].copy()
            O.fn["rdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta, vp(y), 1)
            assert eq_arrays(y, G["r_" + key]), key
        else:
            _, dims, flags, a, b = key.split("_")
            M, N, K = (int(v) for v in dims.split("x"))
            order, ta, tb = flags
            alpha, beta = float(a[1:]), float(b[1:])
            rm = order == "R"
            ar, ac = (M, K) if ta == "N" else (K, M)
            br, bc = (K, N) if tb == "N" else (N, K)
            lda, ldb, ldc = (ac if rm else ar), (bc if rm else br), (N if rm else M)
            A = G[f"A_gemm_{dims}"][: lda * (ar if rm else ac)]
            B = G[f"B_gemm_{dims}"][: ldb * (br if rm else bc)]
            C = G[f"C0_gemm_{dims}"][: ldc * (M if rm else N)].copy()
            O.fn["rdgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb, beta,
                           vp(C), ldc)
            assert eq_arrays(C, G["r_" + key]), key


# --------------------------------------------------------------- properties --
def _exact(xs):
    return sum((Fraction(float(v)) for v in xs), Fraction(0))


@pytest.mark.parametrize("fold", [2, 3, 4])
def test_error_bound_against_exact_rationals(fold):
    """|T - S| < N 2^(40(1-fold)) max|x| + 7 eps |T| (binned_dbbound)."""
    O = oracle()
    for seed in range(40):
        g = dg.rng(seed)
        n = int(g.integers(1, 400))
        x = g.normal(size=n) * np.exp2(g.integers(-60, 60, n).astype(float))
        s = O.rdsum(fold, x)
        t = _exact(x)
        bound = O.dbbound(fold, n, float(np.max(np.abs(x))), float(t))
        assert abs(Fraction(s) - t) <= Fraction(bound), (seed, fold)
        # dot: exact sum of the rounded products
        y = g.normal(size=n)
        d = O.rddot(fold, x, y)
        td = _exact(x * y)
        assert abs(Fraction(d) - td) <= Fraction(O.dbbound(fold, n, float(np.max(np.abs(x * y))), float(td)))


def test_fold3_is_correctly_rounded_on_narrow_range():
    """With every term within 2^30 of the largest, fold 3 keeps > 100 bits, so
    the result is the exact sum correctly rounded (up to the final
    conversion's own rounding, at most 1 ulp here)."""
    O = oracle()
    for seed in range(30):
        x = dg.uniform(2000, 500 + seed) * np.exp2(dg.rng(seed).integers(0, 30, 2000).astype(float))
        s = Fraction(O.rdsum(3, x))
        t = _exact(x)
        ulp = Fraction(math.ulp(float(t)))
        assert abs(s - t) <= ulp


def test_permutation_and_blocking_invariance():
    """verify_daugsum.c: reversals, sorts, shuffles and blocked accumulation
    (binned partials merged with dbdbadd) all give identical bits."""
    O = oracle()
    for kind in ("wide", "ill", "mixed_range", "big"):
        x = _vec(kind, 5000, 3)
        ref = O.rdsum(3, x)
        perms = [x[::-1], np.sort(x), np.sort(x)[::-1], x[np.argsort(np.abs(x))], x[np.argsort(-np.abs(x))]]
        perms += [dg.rng(s).permutation(x) for s in range(5)]
        for p in map(np.ascontiguousarray, perms):
            assert same(O.rdsum(3, p), ref), kind
            assert same(O.rdnrm2(3, p), O.rdnrm2(3, x)), kind
        for blocks in (2, 3, 7, 64):
            acc = np.zeros(6)
            for chunk in np.array_split(x, blocks):
                part = np.zeros(6)
                O.dbdsum(3, chunk, part)
                O.dbdbadd(3, part, acc)
            assert same(O.ddbconv(3, acc), ref), (kind, blocks)
            # sequential accumulation into one binned value
            seq = np.zeros(6)
            for chunk in np.array_split(x, blocks):
                O.dbdsum(3, chunk, seq)
            assert eq_arrays(seq, acc), (kind, blocks)


def test_closed_form_fills():
    """wrap_daugsum.h:232-461 known results."""
    O = oracle()
    small, big = 2.0 ** -27, 2.0 ** 27
    for n in (255, 4095, 8192):
        for scale in (1.0, -1.0, 2.0 ** 500, 2.0 ** -500):
            assert O.rdsum(3, np.full(n, scale)) == n * scale
            v = np.full(n, small * scale)
            v[0] = big * scale
            assert O.rdsum(3, v) == (n - 1) * (scale * small) + scale * big
            v[-1] = -big * scale
            assert O.rdsum(3, v) == (n - 2) * (scale * small)
            assert O.rdasum(3, v) == (n - 2) * (abs(scale) * small) + (abs(scale) * big + abs(scale) * big)
        # sine fill sums to zero exactly: pairs cancel
        h = (n + 1) // 2
        v = np.empty(n)
        v[:h] = np.sin(2.0 * np.pi * (np.arange(h) / n))
        v[0] = 0.0
        for i in range(h, n):
            v[i] = -v[i - n // 2]
        assert abs(O.rdsum(3, v)) <= O.dbbound(3, n, 1.0, 0.0)
        # constant nrm2: sqrt(N (x/s)^2) * s
        x = 0.1
        s = O.dscale(x)
        assert O.rdnrm2(3, np.full(n, x)) == math.sqrt(n * ((x / s) * (x / s))) * s


def test_special_value_classes():
    """dmdsum.ccog:68-78: NaN wins; +inf and -inf give NaN; else the infinity."""
    O = oracle()
    x = np.ones(5000)
    x[10] = np.inf
    assert O.rdsum(3, x) == np.inf
    x[4000] = -np.inf
    assert np.isnan(O.rdsum(3, x))
    x = np.ones(5000)
    x[4999] = -np.inf
    assert O.rdsum(3, x) == -np.inf
    x[0] = np.nan
    assert np.isnan(O.rdsum(3, x))
    assert O.rdnrm2(3, np.array([1.0, -np.inf])) == np.inf
    assert np.isnan(O.rddot(3, np.array([0.0, 1.0]), np.array([np.inf, 1.0])))


def test_index_and_scale_brackets():
    """validate_internal_dindex.c:34-52 and validate_internal_dscale.c:38-56."""
    O = oracle()
    g = dg.rng(5)
    bins = O.lib.orc_dmbins
    bins.restype = ctypes.POINTER(ctypes.c_double)
    bins.argtypes = [ctypes.c_int]
    for e in range(-1021, 1024):
        x = math.ldexp(0.5 + 0.5 * g.uniform(), e)
        i = O.dindex(x)
        xx = x * 2.0 ** -14 if i == 0 else x
        b = bins(i)[0]
        assert b / math.ldexp(0.75, 53) <= 2 * abs(xx)
        assert b / math.ldexp(0.75, 53 - 40) > 2 * abs(xx)
        sc = O.dscale(x)
        m, ex = math.frexp(sc)
        assert m == 0.5 and (ex - 1) % 40 == 0
        assert math.ldexp(0.5, -53 - 40) * sc < x < math.ldexp(0.5, 40 + 3) * sc
    assert O.dindex(0.0) == 51


def test_single_element_round_trip():
    """One term converts back to itself whenever the fold's window spans its
    53 bits: the leading bit may sit up to 40 bits below the top of its bin,
    so folds >= 3 (120 bits) always hold it and fold 2 (80 bits) need not.
    Below about 2^-900 the window reaches the bottom bin (index 51, ulp
    2^-1055) and subnormal tails are cut, as in the reference."""
    O = oracle()
    g = dg.rng(9)
    vals = list(g.normal(size=500) * np.exp2(g.integers(-890, 1020, 500).astype(float)))
    vals += [0.0, 1.7976931348623157e308, -1.7976931348623157e308, 2.0 ** -890]
    for v in vals:
        for fold in (3, 4):
            assert same(O.rdsum(fold, np.array([v])), v), (v, fold)
            assert same(O.ddbconv(fold, O.dbdconv(fold, v)), v)
    # a sum of negative zeros is +0.0 (ddmconv.c:32-34 returns 0.0)
    assert same(O.rdsum(3, -np.zeros(10)), 0.0)


def test_gemm_equals_dot_per_output():
    """wrap_rdgemm.h:16-70: each C element is one binned dot (alpha on A)."""
    O = oracle()
    M, N, K = 9, 7, 300
    A = dg.wide(M * K, 1).reshape(K, M)  # column-major M x K
    B = dg.wide(K * N, 2).reshape(N, K)  # column-major K x N
    C = dg.normal(M * N, 3)
    Cg = C.copy()
    O.fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 2.0, vp(A), M, vp(B), K, 0.5, vp(Cg), M)
    for i in range(M):
        for j in range(N):
            st = O.dbdconv(3, C[j * M + i] * 0.5)
            a = np.ascontiguousarray(A[:, i] * 2.0)
            b = np.ascontiguousarray(B[j, :])
            O.dbddot(3, a, b, st)
            assert same(O.ddbconv(3, st), Cg[j * M + i])
