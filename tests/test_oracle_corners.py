"""Pins the oracle in the two regions where the reference's result depends on
its block schedule rather than on the multiset of terms:

* the infinity path of dmdsum / dmddot / smssum / smsdot
  (src/binnedBLAS/dmdsum.ccog:66-78, dmddot.ccog:68-81): at its first block
  holding an infinite term the reference adds that block's terms one by one
  to its top primary, so a sequential overflow before the infinity turns the
  result into NaN;
* the scaled sum of squares (dmdssq.ccog:60-101, and its complex form
  dmzssq.ccog:59-135 / smcssq.ccog) over values spanning most of the
  exponent range with the largest values in late blocks, at every fold, and
  into an empty accumulator that carries the caller's scale.

Each case is checked against both builds of the compiled reference
(oracle/_ref) when present, and always against tests/golden/corners.npz,
which the compiled reference wrote (tests/golden/make_golden_corners.py).
"""
import ctypes
import os

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle, oracle_s, oracle_z, reference, reference_s, reference_z

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "corners.npz")


def bits64(v):
    return np.array([v], dtype=np.float64).view(np.uint64)[0]


def same(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return bits64(a) == bits64(b)


def same32(a, b):
    a, b = np.float32(a), np.float32(b)
    if np.isnan(a) and np.isnan(b):
        return True
    return a.view(np.uint32) == b.view(np.uint32)


def _golden():
    if not os.path.exists(GOLDEN):
        pytest.skip("tests/golden/corners.npz not written")
    return np.load(GOLDEN)


def corner_cases():
    """(key, op, dtype, kind, n, seed, inc) of every infinity-path case; the
    inputs are regenerated from datagen with these parameters."""
    out = []
    for dt, tag in ((np.float64, "d"), (np.float32, "s")):
        for op in ("sum", "dot"):
            for kind in dg.INF_CORNERS:
                seeds = range(8) if kind == "fuzz" else (1, 2)
                for seed in seeds:
                    for inc in (1, 2):
                        n = 9000 if dt == np.float64 else 3000
                        out.append((f"inf_{tag}_{op}_{kind}_{seed}_{inc}", op, dt, kind, n, seed, inc))
    return out


def make_input(op, dt, kind, n, seed, inc):
    x, y = dg.inf_corner(kind, n, seed, op, dt)
    if inc != 1:
        xs = np.zeros(n * inc, dt)
        xs[::inc] = x
        x = xs
        if y is not None:
            ys = np.zeros(n * inc, dt)
            ys[::inc] = y
            y = ys
    return x, y


def run(lib, op, dt, fold, x, y, inc):
    if dt == np.float64:
        if op == "sum":
            return lib.rdsum(fold, x, inc)
        return lib.rddot(fold, x, y, inc, inc)
    n = len(x) // inc
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    if op == "sum":
        return lib.fn["rssum"](fold, n, p(x), inc)
    return lib.fn["rsdot"](fold, n, p(x), inc, p(y), inc)


CASES = corner_cases()


@pytest.mark.parametrize("key,op,dt,kind,n,seed,inc", CASES, ids=[c[0] for c in CASES])
def test_infinity_path_matches_reference(key, op, dt, kind, n, seed, inc):
    x, y = make_input(op, dt, kind, n, seed, inc)
    orc = oracle() if dt == np.float64 else oracle_s()
    refs = [reference(k) if dt == np.float64 else reference_s(k) for k in ("avx", "sisd")]
    eq = same if dt == np.float64 else same32
    G = _golden()
    for fold in (2, 3, 4):
        e = run(orc, op, dt, fold, x, y, inc)
        for R in refs:
            if R is not None:
                assert eq(e, run(R, op, dt, fold, x, y, inc)), (key, fold)
        assert eq(e, G[f"{key}_f{fold}"][0]), (key, fold, e, G[f"{key}_f{fold}"][0])


def test_infinity_path_cases_reach_both_outcomes():
    """The corner set is not vacuous: it holds results that are NaN only
    because of a sequential overflow (the class of the terms is one
    infinity) and results that stay the infinity."""
    G = _golden()
    nan_one_sign = inf_kept = 0
    for key, op, dt, kind, n, seed, inc in CASES:
        x, y = make_input(op, dt, kind, n, seed, inc)
        with np.errstate(over="ignore", invalid="ignore"):
            t = x[::inc] if op == "sum" else x[::inc] * y[::inc]
        one_sign = not np.isnan(t).any() and not (np.isposinf(t).any() and np.isneginf(t).any())
        r = G[f"{key}_f3"][0]
        if one_sign and np.isnan(r):
            nan_one_sign += 1
        if np.isinf(r):
            inf_kept += 1
    assert nan_one_sign >= 20 and inf_kept >= 20, (nan_one_sign, inf_kept)


def test_infinity_path_accumulate_into_top_state():
    """binnedBLAS_dbdsum into a caller value already at index 0: the
    reference's top primary starts from the caller's, so one big term before
    an infinity overflows it (dmdsum.ccog:66-78)."""
    O = oracle()
    G = _golden()
    for fold in (2, 3, 4):
        Y = np.zeros(2 * fold)
        O.dbdsum(fold, np.array([1.6e308, 1.0]), Y)
        x = dg.uniform(5000, 9)
        x[3000], x[3001] = 1.0e308, -np.inf
        Ye = Y.copy()
        O.dbdsum(fold, x, Ye)
        for k in ("avx", "sisd"):
            R = reference(k)
            if R is not None:
                Yr = Y.copy()
                R.dbdsum(fold, x, Yr)
                assert same(Yr[0], Ye[0]), fold
        assert same(Ye[0], G[f"acc_top_f{fold}"][0]), fold
        assert np.isnan(Ye[0])


def nrm2_range_cases():
    out = []
    for seed in range(6):
        for fold in (2, 3, 4, 8, 13, 20, 26, 27, 30, 40, 52):
            out.append((seed, fold))
    return out


def nrm2_range_input(seed):
    """Values spanning 2^-1000 .. 2^1000 with the largest in late blocks, so
    the reference's running scale rises block by block."""
    g = dg.rng(900 + seed)
    n = 6 * 2048 + 333
    x = dg.uniform(n, 900 + seed) * np.exp2(g.integers(-1000, -600, n).astype(float))
    for b, e in ((1, -300), (2, 0), (3, 400), (4, 800), (5, 1000)):
        x[b * 2048 + int(g.integers(0, 2048))] = 1.25 * 2.0 ** e
    return x


@pytest.mark.parametrize("seed,fold", nrm2_range_cases())
def test_nrm2_wide_range_matches_reference(seed, fold):
    x = nrm2_range_input(seed)
    O = oracle()
    G = _golden()
    e = O.rdnrm2(fold, x)
    Y = np.zeros(2 * fold)
    s = O.dbdssq(fold, x, 0.0, Y)
    for k in ("avx", "sisd"):
        R = reference(k)
        if R is not None:
            assert same(R.rdnrm2(fold, x), e), (k, fold)
            if k == "sisd":
                Yr = np.zeros(2 * fold)
                assert same(R.dbdssq(fold, x, 0.0, Yr), s)
                assert np.array_equal(Yr.view(np.uint64), Y.view(np.uint64)), (fold, Yr, Y)
    assert same(e, G[f"nrm2_{seed}_f{fold}"][0])
    assert np.array_equal(Y.view(np.uint64), G[f"ssq_{seed}_f{fold}"].view(np.uint64))


@pytest.mark.parametrize("fold", [2, 3, 4, 7])
def test_ssq_empty_value_with_scale_matches_reference(fold):
    """dmdssq.ccog:62-101 with Y == 0 and scaleY above dscale(max|x|): the
    scale stays scaleY (for x = [1, 2, 3] and scaleY = dscale(1e60) the
    reference returns scaleY, not dscale(3))."""
    O = oracle()
    G = _golden()
    for j, (x, sy) in enumerate(((np.array([1.0, 2.0, 3.0]), O.dscale(1e60)),
                                 (dg.wide(3000, 530), O.dscale(1e-3)),
                                 (np.zeros(5), O.dscale(1e10)))):
        Y = np.zeros(2 * fold)
        s = O.dbdssq(fold, x, sy, Y)
        assert s == sy or sy < O.dscale(np.max(np.abs(x)))
        for k in ("avx", "sisd"):
            R = reference(k)
            if R is not None:
                Yr = np.zeros(2 * fold)
                assert same(R.dbdssq(fold, x, sy, Yr), s)
                assert np.array_equal(Yr.view(np.uint64), Y.view(np.uint64))
        assert same(s, G[f"ssq0_{j}_f{fold}"][0])
        assert np.array_equal(Y.view(np.uint64), G[f"ssq0_{j}_state_f{fold}"].view(np.uint64))


def table_slots(fold, Y, O):
    """Slots of a binned value whose bin index is inside the bin table
    (index + slot <= 51, binned.h:105), for primaries and carries."""
    I = O.dmindex(Y)
    k = max(0, min(fold, 52 - I))
    return np.r_[0:k, fold:fold + k]


@pytest.mark.parametrize("fold", [28, 30, 40, 52])
def test_ssq_slots_beyond_bin_table_depend_on_the_split(fold):
    """At folds where a value's slots run past bin index 51 (binned.h:105;
    the table repeats its last bin, dmbins.c:23-39), the scaled sum of
    squares puts schedule-dependent content there: one dbdssq call and the
    same terms in two calls merged by dbdbaddsq (dmdmaddsq.c:26-41) differ in
    those slots only, and convert to the same result. The GPU results are
    therefore compared on the slots inside the table at these folds."""
    libs = [oracle()] + [r for r in (reference("avx"), reference("sisd")) if r is not None]
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    differs = 0
    for L in libs:
        for seed in range(6):
            x = nrm2_range_input(seed)
            Y = np.zeros(2 * fold)
            L.dbdssq(fold, x, 0.0, Y)
            A = np.zeros(2 * fold)
            sa = L.dbdssq(fold, np.ascontiguousarray(x[:5000]), 0.0, A)
            B = np.zeros(2 * fold)
            sb = L.dbdssq(fold, np.ascontiguousarray(x[5000:]), 0.0, B)
            L.fn["dbdbaddsq"](fold, sa, vp(A), sb, vp(B))
            keep = table_slots(fold, Y, oracle())
            assert np.array_equal(Y[keep].view(np.uint64), B[keep].view(np.uint64)), (fold, seed)
            assert same(L.ddbconv(fold, Y), L.ddbconv(fold, B))
            differs += not np.array_equal(Y.view(np.uint64), B.view(np.uint64))
    assert differs > 0  # the slots past the table do depend on the split


# ------------------------------------------------------------ amax and NaN --
def amax_nan_cases():
    """(name, x, y or None) with NaN terms at chosen positions: in the middle
    (the max after it counts), last (the result is NaN), several (the last
    one decides), and a NaN product from 0 * inf (damaxm)."""
    g = np.random.default_rng(4242)
    out = []
    for n in (1, 2, 7, 2048, 2049, 100003):
        for pos in sorted({0, n // 3, n - 2, n - 1} - {-1}):
            if pos < 0:
                continue
            x = g.uniform(-4, 4, n)
            x[pos] = np.nan
            if pos + 1 < n:
                x[min(n - 1, pos + 1 + (n - pos) // 2)] = -7.5  # a large term after the NaN
            out.append((f"x_n{n}_nan{pos}", x, None))
    x = g.uniform(-1, 1, 5000)
    x[[10, 2000, 4321]] = np.nan
    x[5] = 99.0  # before the last NaN: dropped
    out.append(("x_three_nans", x, None))
    x = g.uniform(-1, 1, 3000)
    y = g.uniform(-1, 1, 3000)
    x[1500], y[1500] = 0.0, np.inf  # NaN product
    x[1499] = 50.0
    out.append(("xy_zero_times_inf", x, y))
    x = g.uniform(-1, 1, 3000)
    y = g.uniform(-1, 1, 3000)
    x[2999], y[2999] = np.inf, 0.0  # NaN product last
    out.append(("xy_last_nan", x, y))
    return out


AMAX_CASES = amax_nan_cases()


def amax_expected(x, y):
    """The reference's scalar loop restated: max |t| after the last NaN term,
    or NaN when that term is last (damax.ccog:282-300)."""
    with np.errstate(invalid="ignore"):
        t = np.abs(x * y if y is not None else x)
    nan = np.flatnonzero(np.isnan(t))
    if len(nan) == 0:
        return float(max(0.0, t.max()))
    j = nan[-1]
    if j == len(t) - 1:
        return float("nan")
    return float(t[j + 1:].max())


@pytest.mark.parametrize("name,x,y", AMAX_CASES, ids=[c[0] for c in AMAX_CASES])
def test_amax_nan_semantics(name, x, y):
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    O = oracle()
    n = len(x)
    got = O.fn["damaxm"](n, vp(x), 1, vp(y), 1) if y is not None else O.fn["damax"](n, vp(x), 1)
    assert same(got, amax_expected(x, y)), (name, got)
    R = reference("sisd")
    if R is not None:
        r = R.fn["damaxm"](n, vp(x), 1, vp(y), 1) if y is not None else R.fn["damax"](n, vp(x), 1)
        assert same(r, got), (name, r, got)
    # float path: the same positions
    xs = x.astype(np.float32)
    ys = y.astype(np.float32) if y is not None else None
    Os = oracle_s()
    gs = Os.fn["samaxm"](n, vp(xs), 1, vp(ys), 1) if ys is not None else Os.fn["samax"](n, vp(xs), 1)
    assert same32(gs, np.float32(amax_expected(xs, ys))), (name, gs)
    Rs = reference_s("sisd")
    if Rs is not None:
        rs = Rs.fn["samaxm"](n, vp(xs), 1, vp(ys), 1) if ys is not None else Rs.fn["samax"](n, vp(xs), 1)
        assert same32(rs, gs), (name, rs, gs)


# ------------------------------------- nrm2: subnormal squares at deep folds --
def ssq_subnormal_input(dt, seed):
    """Large values in late blocks and tiny ones everywhere whose squares at
    the final scale are subnormal (dmdssq.ccog:60-101 with N_block 2048, or
    smsssq's 512 for float): the reference's binned state then depends on
    its block schedule in the last slot of the bin table."""
    g = np.random.default_rng(7100 + seed)
    O = oracle() if dt == np.float64 else oracle_s()
    nb = 2048 if dt == np.float64 else 512
    if dt == np.float64:
        big = [(1, float(g.integers(-300, 0))), (3, float(g.integers(0, 300))), (5, float(g.integers(300, 1000)))]
        lo, hi = -537.0, -511.0
    else:
        big = [(1, float(g.integers(-40, 0))), (3, float(g.integers(0, 40))), (5, float(g.integers(40, 120)))]
        lo, hi = -75.0, -63.0
    n = 6 * nb + int(g.integers(0, nb))
    M = max(1.5 * 2.0 ** e for _, e in big)
    S = float(O.dscale(M)) if dt == np.float64 else float(O.fn["sscale"](np.float32(M)))
    x = (S * g.uniform(1, 2, n) * np.exp2(g.uniform(lo, hi, n))) * g.choice([-1.0, 1.0], n)
    for b, e in big:
        x[b * nb + int(g.integers(0, nb))] = 1.5 * 2.0 ** e
    return x.astype(dt)


SSQ_SUB_CASES = [(np.float64, seed, fold) for seed in range(4) for fold in (27, 30, 40, 52)] + \
                [(np.float32, seed, fold) for seed in range(4) for fold in (8, 12, 16, 21)]


@pytest.mark.parametrize("dt,seed,fold", SSQ_SUB_CASES,
                         ids=[f"{'d' if c[0] == np.float64 else 's'}{c[1]}_f{c[2]}" for c in SSQ_SUB_CASES])
def test_ssq_subnormal_squares_match_reference(dt, seed, fold):
    x = ssq_subnormal_input(dt, seed)
    d = dt == np.float64
    O = oracle() if d else oracle_s()
    Y = np.zeros(2 * fold, dtype=dt)
    s = O.dbdssq(fold, x, 0.0, Y) if d else O.sbsssq(fold, x, 0.0, Y)
    eq = same if d else same32
    R = reference("sisd") if d else reference_s("sisd")
    if R is not None:
        Yr = np.zeros(2 * fold, dtype=dt)
        sr = R.dbdssq(fold, x, 0.0, Yr) if d else R.sbsssq(fold, x, 0.0, Yr)
        assert eq(sr, s)
        u = np.uint64 if d else np.uint32
        assert np.array_equal(Yr.view(u), Y.view(u)), (seed, fold)
        assert eq(R.rdnrm2(fold, x) if d else R.rsnrm2(fold, x), O.rdnrm2(fold, x) if d else O.rsnrm2(fold, x))


def zssq_subnormal_input(prec, seed, blocks):
    """The complex counterpart (dmzssq.ccog:59-135 with N_block 1024, or
    smcssq's 256 for float), interleaved (re, im): large components in the
    given blocks, tiny ones everywhere whose squares at the final scale are
    subnormal. The reference renormalises before every second block only, so
    a large value in an odd block rescales an unrenormalised accumulator."""
    g = np.random.default_rng(7300 + seed)
    d = prec == "z"
    O = oracle() if d else oracle_s()
    nb = 1024 if d else 256
    if d:
        exps = [float(g.integers(-300, 0)), float(g.integers(0, 300)), float(g.integers(300, 1000))]
        lo, hi = -537.0, -511.0
    else:
        exps = [float(g.integers(-40, 0)), float(g.integers(0, 40)), float(g.integers(40, 120))]
        lo, hi = -75.0, -63.0
    n = 6 * nb + int(g.integers(0, nb))
    M = max(1.5 * 2.0 ** e for e in exps)
    S = float(O.dscale(M)) if d else float(O.fn["sscale"](np.float32(M)))
    x = (S * g.uniform(1, 2, 2 * n) * np.exp2(g.uniform(lo, hi, 2 * n))) * g.choice([-1.0, 1.0], 2 * n)
    for b, e in zip(blocks, exps):
        x[2 * (b * nb + int(g.integers(0, nb))) + int(g.integers(0, 2))] = 1.5 * 2.0 ** e
    return x.astype(np.float64 if d else np.float32)


ZSSQ_SUB_CASES = [(prec, seed, blocks, fold) for prec, folds in (("z", (27, 40, 52)), ("c", (9, 12, 21)))
                  for seed in range(3) for blocks in ((1, 2, 5), (1, 3, 5)) for fold in folds]
ZSSQ_SUB_IDS = [f"{c[0]}{c[1]}_b{''.join(map(str, c[2]))}_f{c[3]}" for c in ZSSQ_SUB_CASES]


@pytest.mark.parametrize("prec,seed,blocks,fold", ZSSQ_SUB_CASES, ids=ZSSQ_SUB_IDS)
def test_zssq_subnormal_squares_match_reference(prec, seed, blocks, fold):
    x = zssq_subnormal_input(prec, seed, blocks)
    Z = oracle_z(prec)
    u = np.uint64 if prec == "z" else np.uint32
    Y = np.zeros(2 * fold, dtype=x.dtype)
    s = Z.bssq(fold, x, 0.0, Y)
    R = reference_z(prec, "sisd")
    if R is not None:
        Yr = np.zeros(2 * fold, dtype=x.dtype)
        sr = R.bssq(fold, x, 0.0, Yr)
        assert s == sr
        assert np.array_equal(Yr.view(u), Y.view(u)), (seed, blocks, fold)
        assert np.array_equal(np.array([R.rnrm2(fold, x)], x.dtype).view(u), np.array([Z.rnrm2(fold, x)], x.dtype).view(u))


def test_zssq_subnormal_inputs_reach_the_schedule_dependent_slots():
    """The complex family is not vacuous: the same terms in two calls merged
    by the scaled add give another state for most cases."""
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    for prec in ("z", "c"):
        Z = oracle_z(prec)
        O = oracle() if prec == "z" else oracle_s()
        u = np.uint64 if prec == "z" else np.uint32
        differs = 0
        for _, seed, blocks, fold in [c for c in ZSSQ_SUB_CASES if c[0] == prec]:
            x = zssq_subnormal_input(prec, seed, blocks)
            Y = np.zeros(2 * fold, dtype=x.dtype)
            Z.bssq(fold, x, 0.0, Y)
            h = (len(x) // 4) * 2
            A, B = np.zeros(2 * fold, dtype=x.dtype), np.zeros(2 * fold, dtype=x.dtype)
            sa = Z.bssq(fold, np.ascontiguousarray(x[:h]), 0.0, A)
            sb = Z.bssq(fold, np.ascontiguousarray(x[h:]), 0.0, B)
            if prec == "z":
                O.fn["dbdbaddsq"](fold, sa, vp(A), sb, vp(B))
            else:
                O.fn["sbsbaddsq"](fold, np.float32(sa), vp(A), np.float32(sb), vp(B))
            differs += not np.array_equal(Y.view(u), B.view(u))
        assert differs >= 6, (prec, differs)


def test_ssq_subnormal_inputs_reach_the_schedule_dependent_slot():
    """The family above is not vacuous: split into two calls merged by
    dbdbaddsq, the reference's state differs inside the bin table."""
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    R = reference("sisd") or oracle()
    O = oracle()
    differs = 0
    for seed in range(4):
        for fold in (30, 40):
            x = ssq_subnormal_input(np.float64, seed)
            Y = np.zeros(2 * fold)
            R.dbdssq(fold, x, 0.0, Y)
            A, B = np.zeros(2 * fold), np.zeros(2 * fold)
            sa = R.dbdssq(fold, np.ascontiguousarray(x[:3000]), 0.0, A)
            sb = R.dbdssq(fold, np.ascontiguousarray(x[3000:]), 0.0, B)
            R.fn["dbdbaddsq"](fold, sa, vp(A), sb, vp(B))
            keep = table_slots(fold, Y, O)
            differs += not np.array_equal(Y[keep].view(np.uint64), B[keep].view(np.uint64))
    assert differs > 0
