"""Pins the single-precision oracle (oracle/oracle_s.c) before it is trusted
as the reference for the FP32 kernels.

1. Bit-equality with the compiled reference (oracle/_ref: the reference's own
   s-path C sources, AVX and scalar builds) over random, ill-conditioned,
   extreme and non-finite inputs, N straddling the 512-element block edge,
   strides, folds 2..21, every primitive, gemv and gemm layouts.
2. Properties the mathematics fixes: the reference's bound (binned_sbbound)
   against exact rational sums; permutation and blocking invariance; closed
   forms; single-term round trips.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle_s, reference_s


def same(a, b):
    a, b = np.float32(a), np.float32(b)
    if np.isnan(a) and np.isnan(b):
        return True
    return np.array([a]).view(np.uint32)[0] == np.array([b]).view(np.uint32)[0]


def eq_arrays(a, b):
    a, b = np.asarray(a, dtype=np.float32), np.asarray(b, dtype=np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint32) == b.view(np.uint32))))


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def _refs():
    out = [r for r in (reference_s("avx"), reference_s("sisd")) if r is not None]
    if not out:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return out


INPUTS = ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "constant", "nan", "pinf", "ninf",
          "both_inf", "inf_nan"]


def _vec(kind, n, seed):
    if kind == "uniform":
        return dg.uniform(n, seed).astype(np.float32)
    if kind == "wide":
        return dg.wide_s(n, seed)
    if kind == "ill":
        return dg.ill_conditioned(n, seed).astype(np.float32)
    if n < 2 and kind in ("both_inf", "inf_nan"):
        return dg.special_s("nan", n, seed)
    return dg.special_s(kind, n, seed)


@pytest.mark.parametrize("kind", INPUTS)
@pytest.mark.parametrize("fold", [2, 3, 4, 5, 8, 13, 21])
def test_blas1_matches_compiled_reference(kind, fold):
    O = oracle_s()
    for R in _refs():
        for n in (1, 2, 3, 511, 512, 513, 1025, 3000):
            x, y = _vec(kind, n, 10 + n), dg.wide_s(n, 20 + n)
            for op in ("rssum", "rsasum", "rsnrm2"):
                assert same(getattr(O, op)(fold, x), getattr(R, op)(fold, x)), (kind, fold, n, op)
            assert same(O.rsdot(fold, x, y), R.rsdot(fold, x, y)), (kind, fold, n)
            for st_op in ("sbssum", "sbsasum"):
                a, b = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
                getattr(O, st_op)(fold, x, a)
                getattr(R, st_op)(fold, x, b)
                if not np.isnan(a[0]):
                    assert eq_arrays(a, b), (kind, fold, n, st_op)
            a, b = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
            sa, sb = O.sbsssq(fold, x, 0.0, a), R.sbsssq(fold, x, 0.0, b)
            assert same(sa, sb)
            if not np.isnan(a[0]) and not eq_arrays(a, b):
                # The reference's AVX build differs from its own scalar build
                # when rescaled ssq primaries are subnormal (deep folds); the
                # oracle follows the scalar build, which must agree exactly.
                c = np.zeros(2 * fold, np.float32)
                S = reference_s("sisd")
                assert R is not S and S is not None
                S.sbsssq(fold, x, 0.0, c)
                assert eq_arrays(a, c) and not eq_arrays(b, c), (kind, fold, n)
                assert np.all(np.abs(a[~(a.view(np.uint32) == b.view(np.uint32))]) < np.float32(1.2e-38))


@pytest.mark.parametrize("inc", [2, -1, -3])
def test_strided_matches_reference(inc):
    O = oracle_s()
    n = 1501
    base = dg.wide_s(n * abs(inc), 5)
    ptr = ctypes.c_void_p(base.ctypes.data + (0 if inc > 0 else (n - 1) * abs(inc) * 4))
    for R in _refs():
        for op in ("rssum", "rsasum", "rsnrm2"):
            assert same(O.fn[op](3, n, ptr, inc), R.fn[op](3, n, ptr, inc))
        assert same(O.fn["rsdot"](3, n, ptr, inc, ptr, inc), R.fn["rsdot"](3, n, ptr, inc, ptr, inc))


def test_accumulate_chain_matches_reference():
    O = oracle_s()
    for R in _refs():
        for fold in (2, 3, 4):
            a, b = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
            sa = sb = 0.0
            qa, qb = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
            for k, x in enumerate([dg.wide_s(900, 1), (dg.wide_s(5000, 2) * np.float32(1e10)),
                                   (dg.uniform(3, 3) * 1e-10).astype(np.float32), dg.special_s("big", 100, 4)]):
                O.sbssum(fold, x, a)
                R.sbssum(fold, x, b)
                assert eq_arrays(a, b), (fold, k)
                sa, sb = O.sbsssq(fold, x, sa, qa), R.sbsssq(fold, x, sb, qb)
                assert same(sa, sb) and eq_arrays(qa, qb), (fold, k)


def test_primitives_match_reference():
    O = oracle_s()
    g = dg.rng(77)
    vals = [np.float32(v) for v in np.ldexp(1.0, np.arange(-149, 128, 1))]
    vals += [np.float32(v) for v in (0.0, -0.0, np.inf, -np.inf, np.nan, 1.4e-45, 3.4028234e38, 1.1754942e-38)]
    vals += list(dg.special_s("tiny", 200, 3))
    for R in _refs():
        for v in vals:
            assert O.fn["sindex"](v) == R.fn["sindex"](v), v
            assert same(O.fn["sscale"](v), R.fn["sscale"](v)), v
        for fold in (2, 3, 4, 7, 21):
            for _ in range(200):
                xs = (g.normal(size=5) * np.exp2(g.integers(-150, 125))).astype(np.float32)
                ys = (g.normal(size=5) * np.exp2(g.integers(-150, 125))).astype(np.float32)
                X1, X2, Y1, Y2 = (np.zeros(2 * fold, np.float32) for _ in range(4))
                for v in xs:
                    O.sbsadd(fold, v, X1)
                    R.sbsadd(fold, v, X2)
                for v in ys:
                    O.sbsadd(fold, v, Y1)
                    R.sbsadd(fold, v, Y2)
                assert eq_arrays(X1, X2) and eq_arrays(Y1, Y2)
                assert same(O.ssbconv(fold, X1), R.ssbconv(fold, X2))
                O.sbsbadd(fold, X1, Y1)
                R.sbsbadd(fold, X2, Y2)
                assert eq_arrays(Y1, Y2)
                s1, s2 = O.fn["sscale"](abs(xs[0])), O.fn["sscale"](abs(ys[0]))
                A1, A2 = Y1.copy(), Y2.copy()
                assert same(O.fn["sbsbaddsq"](fold, s1, vp(X1), s2, vp(A1)),
                            R.fn["sbsbaddsq"](fold, s1, vp(X2), s2, vp(A2)))
                assert eq_arrays(A1, A2)
            assert same(O.fn["sbbound"](fold, 1000, 3.5, 17.0), R.fn["sbbound"](fold, 1000, 3.5, 17.0))


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_gemv_matches_reference(order, trans):
    O = oracle_s()
    M, N = 131, 97
    lda = (M if order == "C" else N) + 2
    A = dg.wide_s(lda * max(M, N), 31)
    x = dg.normal(max(M, N), 32).astype(np.float32)
    y0 = dg.normal(max(M, N), 33).astype(np.float32)
    for R in _refs():
        for fold in (2, 3):
            for alpha, beta in ((1.0, 0.0), (2.0, 0.5), (0.0, 1.0), (0.5, 1.0)):
                ya, yb = y0.copy(), y0.copy()
                for L, y in ((O, ya), (R, yb)):
                    L.fn["rsgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                   vp(y), 1)
                assert eq_arrays(ya, yb), (fold, alpha, beta)


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_matches_reference(order, ta, tb):
    O = oracle_s()
    M, N, K = 37, 23, 1100
    A = dg.wide_s(K * max(M, N) + 64, 41)
    B = dg.wide_s(K * max(M, N) + 64, 42)
    C0 = dg.normal(max(M, N) ** 2 + 64, 43).astype(np.float32)
    rm = order == "R"
    lda = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
    ldb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
    ldc = N if rm else M
    for R in _refs():
        for alpha, beta in ((1.0, 0.0), (3.0, 0.5), (1.0, 1.0), (0.0, 2.0)):
            ca, cb = C0.copy(), C0.copy()
            for L, C in ((O, ca), (R, cb)):
                L.fn["rsgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb,
                               beta, vp(C), ldc)
            assert eq_arrays(ca, cb), (alpha, beta)


# ------------------------------------------------------------ properties --
@pytest.mark.parametrize("fold", [2, 3, 4])
def test_bound_against_exact_sums(fold):
    """|rssum - exact| <= binned_sbbound(fold, N, max|x|, exact) (the
    reference's error bound), with the exact sum in rationals."""
    O = oracle_s()
    for seed in range(20):
        n = 50 + 37 * seed
        x = dg.wide_s(n, 300 + seed)
        exact = sum(Fraction(float(v)) for v in x)
        got = O.rssum(fold, x)
        bound = O.fn["sbbound"](fold, n, float(np.max(np.abs(x))), float(exact))
        assert abs(Fraction(float(got)) - exact) <= Fraction(float(bound)), (fold, seed)


def test_order_and_blocking_invariance():
    O = oracle_s()
    x = dg.wide_s(20000, 9)
    for fold in (2, 3, 4):
        ref = O.rssum(fold, x)
        for seed in (1, 2, 3):
            assert same(O.rssum(fold, dg.rng(seed).permutation(x)), ref)
        Y = np.zeros(2 * fold, np.float32)
        for lo in range(0, 20000, 777):
            O.sbssum(fold, np.ascontiguousarray(x[lo:lo + 777]), Y)
        assert same(O.ssbconv(fold, Y), ref)


def test_closed_forms_and_round_trips():
    O = oracle_s()
    for fold in (2, 3):
        assert O.rssum(fold, np.full(4096, 0.5, np.float32)) == 2048.0
        v = dg.wide_s(1000, 4)
        assert O.rssum(fold, np.concatenate([v, -v])) == 0.0
    # one term round-trips exactly once three 13-bit bins cover its 24 bits
    for fold in (3, 4):
        for t in list(dg.wide_s(200, 5)) + [np.float32(-3.0e38), np.float32(1e-30)]:
            Y = np.zeros(2 * fold, np.float32)
            O.fn["sbsconv"](fold, t, vp(Y))
            assert same(O.ssbconv(fold, Y), t), (fold, t)


# ---------------------------------------------------------------- golden --
def _golden_sc():
    import os

    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sc.npz")
    return np.load(p)


def test_oracle_matches_golden_fixtures():
    """tests/golden/sc.npz holds the compiled reference's single-precision
    outputs (tests/golden/make_golden_sc.py); the oracle reproduces them
    without the reference library."""
    O = oracle_s()
    G = _golden_sc()
    keys = [k for k in G.files if k.startswith("s_") and k.endswith("_res")]
    assert keys
    for k in keys:
        parts = k.split("_")
        kind, n, fold = "_".join(parts[1:-3]), parts[-3], int(parts[-2])
        x, y = G[f"s_{kind}_{n}_x"], G[f"s_{kind}_{n}_y"]
        got = np.array([O.rssum(fold, x), O.rsasum(fold, x), O.rsnrm2(fold, x), O.rsdot(fold, x, y)], np.float32)
        assert eq_arrays(got, G[k]), k
        st = np.zeros(2 * fold, np.float32)
        O.sbssum(fold, x, st)
        if not np.isnan(st[0]):
            assert eq_arrays(st, G[f"s_{kind}_{n}_{fold}_sbssum"]), k
    A, B, C0 = G["s_mat_A"], G["s_mat_B"], G["s_mat_C"]
    M, N, K = 37, 29, 300
    for order in "CR":
        for tr in "NT":
            y = C0.copy()
            lda = (M if order == "C" else N) + 1
            O.fn["rsgemv"](3, order.encode(), tr.encode(), M, N, 1.5, vp(A), lda, vp(B), 1, 0.5, vp(y), 1)
            assert eq_arrays(y, G[f"s_gemv_{order}{tr}"]), (order, tr)
            for tb in "NT":
                rm = order == "R"
                la = (K if tr == "N" else M) if rm else (M if tr == "N" else K)
                lb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
                c = C0.copy()
                O.fn["rsgemm"](3, order.encode(), tr.encode(), tb.encode(), M, N, K, 0.75, vp(A), la, vp(B), lb, 0.5,
                               vp(c), N if rm else M)
                assert eq_arrays(c, G[f"s_gemm_{order}{tr}{tb}"]), (order, tr, tb)
