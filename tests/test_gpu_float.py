"""GPU parity of the single-precision path (float inputs, float_binned states
of 2*fold floats, folds 2..21) against the oracle's single-precision calls.

Bar: bit-identical floats (NaN compared by class). Calls go through the C ABI
of libreproblas_hip.so with host (staged) and device (in place) operands.
Reference behaviour: src/reproBLAS/rs{sum,asum,nrm2,dot}.c, the sbs* drivers
(src/binnedBLAS/sbs*.c, smssum.ccog and siblings), rsgemv.c / rsgemm.c.
"""
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle_s

pytestmark = pytest.mark.gpu


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


def orc_call(op, fold, x, y=None):
    O = oracle_s()
    if op == "sum":
        return O.rssum(fold, x)
    if op == "asum":
        return O.rsasum(fold, x)
    if op == "nrm2":
        return O.rsnrm2(fold, x)
    return O.rsdot(fold, x, y)


def gpu_call(rb, op, fold, x, y=None):
    if op == "sum":
        return rb.rssum(x, fold)
    if op == "asum":
        return rb.rsasum(x, fold)
    if op == "nrm2":
        return rb.rsnrm2(x, fold)
    return rb.rsdot(x, y, fold)


OPS = ["sum", "asum", "nrm2", "dot"]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("fold", [2, 3, 4, 8])
@pytest.mark.parametrize("n", [1, 7, 511, 512, 513, 4095, 65537, 1 << 20])
def test_random_host_operands(gpu, op, fold, n):
    import reproblas_amd as rb

    x = dg.wide_s(n, 11 + n)
    y = dg.wide_s(n, 12 + n)
    r = gpu_call(rb, op, fold, x, y)
    e = orc_call(op, fold, x, y)
    assert same(r, e), (op, fold, n, r, e)


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("dist", ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "neg_zeros",
                                  "constant"])
def test_distributions_device_operands(gpu, op, dist):
    import torch

    import reproblas_amd as rb

    n = 100003
    if dist == "uniform":
        x, y = dg.uniform(n, 1).astype(np.float32), dg.uniform(n, 2).astype(np.float32)
    elif dist == "wide":
        x, y = dg.wide_s(n, 1), dg.wide_s(n, 2)
    elif dist == "ill":
        x, y = dg.ill_conditioned(n, 1).astype(np.float32), dg.uniform(n, 2).astype(np.float32)
    else:
        x, y = dg.special_s(dist, n, 3), dg.special_s(dist, n, 4)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    for fold in (2, 3, 4):
        r = gpu_call(rb, op, fold, tx, ty)
        e = orc_call(op, fold, x, y)
        assert same(r, e), (op, dist, fold, r, e)


# special fills x lengths; both_inf and inf_nan need two special members
SPECIAL_CASES = [(k, n) for k in ("nan", "pinf", "ninf", "both_inf", "inf_nan") for n in (1, 300, 70001)
                 if n >= 2 or k not in ("both_inf", "inf_nan")]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("kind,n", SPECIAL_CASES)
def test_special_values(gpu, op, kind, n):
    import reproblas_amd as rb

    x = dg.special_s(kind, n, 5)
    y = dg.uniform(n, 6).astype(np.float32)
    r = gpu_call(rb, op, 3, x, y)
    e = orc_call(op, 3, x, y)
    assert same(r, e), (op, kind, n, r, e)


@pytest.mark.parametrize("op", ["sum", "asum", "nrm2"])
@pytest.mark.parametrize("inc", [2, 3, -1, -2])
def test_strides(gpu, op, inc):
    import reproblas_amd as rb

    n = 20001
    base = dg.wide_s(n * abs(inc), 7)
    O = oracle_s()
    name = {"sum": "rssum", "asum": "rsasum", "nrm2": "rsnrm2"}[op]
    # negative increments: the pointer is the lowest-addressed element, which
    # the reference indexes last (X[i*inc] walks backwards from the far end)
    ptr = ctypes.c_void_p(base.ctypes.data + (0 if inc > 0 else (n - 1) * abs(inc) * 4))
    e = O.fn[name](3, n, ptr, inc)
    L = rb.lib()
    rb._bind_stream()
    r = getattr(L, "reproBLAS_" + name)(3, n, ptr, inc)
    rb._check()
    assert same(r, e), (op, inc, r, e)


def test_dot_strided(gpu):
    import reproblas_amd as rb

    n = 30001
    x = dg.wide_s(2 * n, 8)
    y = dg.wide_s(3 * n, 9)
    e = oracle_s().fn["rsdot"](3, n, vp(x), 2, vp(y), 3)
    r = rb.rsdot(x, y, 3, 2, 3, n=n)
    assert same(r, e)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ill_conditioned_shuffles_identical(gpu, seed):
    import reproblas_amd as rb

    n = 1 << 20
    base = rb.rssum(dg.ill_conditioned(n, 0).astype(np.float32))
    xs = dg.ill_conditioned(n, seed).astype(np.float32)
    r = rb.rssum(xs)
    assert same(r, oracle_s().rssum(3, xs))
    # the float values differ from seed 0's after rounding, so only the
    # permutation of one rounded vector must give the same bits
    p = dg.rng(seed).permutation(xs)
    assert same(rb.rssum(p), r)
    del base


@pytest.mark.parametrize("op", ["sum", "asum", "dot"])
def test_accumulate_into_binned(gpu, op):
    """sbs* accumulate into an existing float_binned value (host and device Y)."""
    import torch

    import reproblas_amd as rb

    O = oracle_s()
    for fold in (2, 3, 4):
        x1, x2 = dg.wide_s(5000, 21), dg.wide_s(7000, 22) * np.float32(1e6)
        y1, y2 = dg.wide_s(5000, 23), dg.wide_s(7000, 24)
        Yh = np.zeros(2 * fold, np.float32)
        Ye = np.zeros(2 * fold, np.float32)
        for xa, ya in ((x1, y1), (x2, y2)):
            if op == "sum":
                rb.sbssum(Yh, xa, fold)
                O.sbssum(fold, xa, Ye)
            elif op == "asum":
                rb.sbsasum(Yh, xa, fold)
                O.sbsasum(fold, xa, Ye)
            else:
                rb.sbsdot(Yh, xa, ya, fold)
                O.sbsdot(fold, xa, ya, Ye)
        assert eq_arrays(Yh, Ye), (op, fold, Yh, Ye)
        Yd = torch.zeros(2 * fold, dtype=torch.float32, device=gpu)
        for xa, ya in ((x1, y1), (x2, y2)):
            tx, ty = torch.from_numpy(xa).to(gpu), torch.from_numpy(ya).to(gpu)
            if op == "sum":
                rb.sbssum(Yd, tx, fold)
            elif op == "asum":
                rb.sbsasum(Yd, tx, fold)
            else:
                rb.sbsdot(Yd, tx, ty, fold)
        assert eq_arrays(Yd.cpu().numpy(), Ye)


def test_ssq_accumulate_scale(gpu):
    import reproblas_amd as rb

    O = oracle_s()
    for fold in (2, 3, 4):
        Yh, Ye = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
        sh = se = 0.0
        for k, xa in enumerate((dg.wide_s(3000, 31), dg.wide_s(4000, 32) * np.float32(1e15),
                                dg.wide_s(100, 33) * np.float32(1e-15))):
            sh = rb.sbsssq(Yh, xa, sh, fold)
            se = O.sbsssq(fold, xa, se, Ye)
            assert same(sh, se), (fold, k, sh, se)
            assert eq_arrays(Yh, Ye), (fold, k, Yh, Ye)


def test_empty_and_bad_fold(gpu):
    import reproblas_amd as rb

    x = np.zeros(0, np.float32)
    assert rb.rssum(x) == 0.0
    assert rb.rsnrm2(x) == 0.0
    assert rb.lib().reproBLAS_rssum(3, -5, None, 1) == 0.0
    with pytest.raises(rb.ReproBLASError):
        rb.rssum(np.ones(10, np.float32), fold=1)
    with pytest.raises(rb.ReproBLASError):
        rb.rssum(np.ones(10, np.float32), fold=22)
    with pytest.raises(TypeError):
        rb.rssum(np.ones(10))


def test_device_result_api(gpu):
    """reproblas_hip_rs*_device / sbs*_device: async, results in device memory."""
    import torch

    import reproblas_amd as rb

    n = 1 << 18
    x, y = dg.wide_s(n, 41), dg.wide_s(n, 42)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    out = torch.zeros(4, dtype=torch.float32, device=gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = rb.lib()
    p = lambda t, off=0: ctypes.c_void_p(t.data_ptr() + 4 * off)
    assert L.reproblas_hip_rssum_device(3, n, p(tx), 1, p(out, 0), s) == 0
    assert L.reproblas_hip_rsasum_device(3, n, p(tx), 1, p(out, 1), s) == 0
    assert L.reproblas_hip_rsnrm2_device(3, n, p(tx), 1, p(out, 2), s) == 0
    assert L.reproblas_hip_rsdot_device(3, n, p(tx), 1, p(ty), 1, p(out, 3), s) == 0
    got = out.cpu().numpy()
    O = oracle_s()
    exp = [O.rssum(3, x), O.rsasum(3, x), O.rsnrm2(3, x), O.rsdot(3, x, y)]
    for g_, e_ in zip(got, exp):
        assert same(g_, e_)
    # binned-state device calls
    for fold in (2, 3):
        Y = torch.zeros(2 * fold, dtype=torch.float32, device=gpu)
        Z = torch.zeros(2 * fold, dtype=torch.float32, device=gpu)
        Q = torch.zeros(2 * fold, dtype=torch.float32, device=gpu)
        sc = torch.zeros(1, dtype=torch.float32, device=gpu)
        assert L.reproblas_hip_sbssum_device(fold, n, p(tx), 1, p(Y), s) == 0
        assert L.reproblas_hip_sbsdot_device(fold, n, p(tx), 1, p(ty), 1, p(Z), s) == 0
        assert L.reproblas_hip_sbsssq_device(fold, n, p(tx), 1, p(sc), p(Q), s) == 0
        Ye, Ze, Qe = (np.zeros(2 * fold, np.float32) for _ in range(3))
        O.sbssum(fold, x, Ye)
        O.sbsdot(fold, x, y, Ze)
        se = O.sbsssq(fold, x, 0.0, Qe)
        assert eq_arrays(Y.cpu().numpy(), Ye)
        assert eq_arrays(Z.cpu().numpy(), Ze)
        assert eq_arrays(Q.cpu().numpy(), Qe)
        assert same(sc.cpu().numpy()[0], se)


def test_samax(gpu):
    import reproblas_amd as rb

    x = dg.wide_s(12345, 51)
    y = dg.wide_s(12345, 52)
    O = oracle_s()
    assert same(rb.samax(x), O.fn["samax"](len(x), vp(x), 1))
    assert same(rb.samaxm(x, y), O.fn["samaxm"](len(x), vp(x), 1, vp(y), 1))


def test_binned_vector_add_and_conversions(gpu):
    """binned_sbsbaddv on device vectors of float_binned, against sbsbadd per
    element in the oracle; sbsconv / ssbconv host round trips."""
    import torch

    import reproblas_amd as rb

    O = oracle_s()
    for fold in (2, 3, 5):
        n = 1000
        X = np.zeros((n, 2 * fold), np.float32)
        Y = np.zeros((n, 2 * fold), np.float32)
        vals = dg.wide_s(3 * n, 61)
        for i in range(n):
            O.sbsadd(fold, vals[i], X[i])
            O.sbsadd(fold, vals[n + i], Y[i])
            O.sbsadd(fold, vals[2 * n + i] * np.float32(1e8), Y[i])
        tX, tY = torch.from_numpy(X).to(gpu), torch.from_numpy(Y).to(gpu)
        rb.sbsbaddv(tX, tY, n, fold)
        for i in range(n):
            O.sbsbadd(fold, X[i], Y[i])
        assert eq_arrays(tY.cpu().numpy(), Y), fold
        for v in vals[:50]:
            B = rb.sbsconv(float(v), fold)
            assert same(rb.ssbconv(B, fold), O.ssbconv(fold, B))


@pytest.mark.parametrize("fold", [9, 13, 21])
def test_high_folds(gpu, fold):
    """Folds above 8 run the general-capacity kernels; 21 is float_binned's maximum."""
    import torch

    import reproblas_amd as rb

    O = oracle_s()
    for n in (1, 777, 70001):
        x, y = dg.wide_s(n, 300 + n), dg.wide_s(n, 400 + n)
        for op in OPS:
            assert same(gpu_call(rb, op, fold, x, y), orc_call(op, fold, x, y)), (fold, n, op)
    x = dg.special_s("mixed_range", 5000, 9)
    assert same(rb.rssum(x, fold), O.rssum(fold, x))
    Y, Ye = np.zeros(2 * fold, np.float32), np.zeros(2 * fold, np.float32)
    rb.sbssum(Y, x, fold)
    O.sbssum(fold, x, Ye)
    assert eq_arrays(Y, Ye)
    tx = torch.from_numpy(x).to(gpu)
    assert same(rb.rsnrm2(tx, fold), O.rsnrm2(fold, x))


# ---------------------------------------------------------------- gemv/gemm --
@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (255, 512), (300, 77), (1000, 2049)])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (2.0, 0.5), (0.5, 1.0), (0.0, 2.0)])
def test_rsgemv(gpu, order, trans, shape, ab):
    import reproblas_amd as rb

    M, N = shape
    alpha, beta = ab
    lda = (M if order == "C" else N) + 3
    A = dg.normal(lda * (N if order == "C" else M), 61).astype(np.float32)
    lx = N if trans == "N" else M
    ly = M if trans == "N" else N
    x = dg.normal(lx, 62).astype(np.float32)
    y0 = dg.normal(ly, 63).astype(np.float32)
    for fold in (2, 3, 4):
        ye = y0.copy()
        oracle_s().fn["rsgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                vp(ye), 1)
        yg = y0.copy()
        rb.lib().reproBLAS_rsgemv(fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                  vp(yg), 1)
        rb._check()
        assert eq_arrays(yg, ye), (order, trans, shape, ab, fold)


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (64, 64, 64), (96, 80, 72), (33, 65, 130), (128, 40, 2100)])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (2.0, 0.5), (0.5, 1.0)])
def test_rsgemm(gpu, order, ta, tb, mnk, ab):
    import reproblas_amd as rb

    M, N, K = mnk
    alpha, beta = ab
    rm = order == "R"
    lda = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
    ldb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
    ldc = N if rm else M
    A = dg.wide_s(M * K + 16, 71)
    B = dg.wide_s(K * N + 16, 72)
    C0 = dg.normal(M * N + 16, 73).astype(np.float32)
    ce = C0.copy()
    oracle_s().fn["rsgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb,
                            beta, vp(ce), ldc)
    cg = C0.copy()
    rb.lib().reproBLAS_rsgemm(3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb,
                              beta, vp(cg), ldc)
    rb._check()
    assert eq_arrays(cg, ce), (order, ta, tb, mnk, ab)


@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_gemm_tilings_identical(gpu, tiling):
    """Every gemm tiling gives the oracle's bits on device operands."""
    import torch

    import reproblas_amd as rb

    M, N, K = 200, 136, 700
    A = dg.wide_s(M * K, 81).reshape(K, M)
    B = dg.wide_s(K * N, 82).reshape(N, K)
    C = np.zeros((N, M), np.float32)
    ce = C.copy()
    oracle_s().fn["rsgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.0, vp(ce), M)
    old = rb.get_gemm_tiling()
    try:
        rb.set_gemm_tiling(tiling)
        tA, tB, tC = (torch.from_numpy(a).to(gpu) for a in (A, B, C))
        rb.rsgemm("N", "N", 1.0, tA, tB, 0.0, tC)
        assert eq_arrays(tC.cpu().numpy(), ce)
    finally:
        rb.set_gemm_tiling(old)


def test_gemm_special_values(gpu):
    import reproblas_amd as rb

    M, N, K = 40, 30, 300
    A = dg.wide_s(M * K, 91)
    B = dg.wide_s(K * N, 92)
    A[5] = np.inf
    A[77] = -np.inf
    B[100] = np.nan
    B[K * 3 + 7] = np.float32(3e38)
    C0 = np.zeros(M * N, np.float32)
    ce, cg = C0.copy(), C0.copy()
    oracle_s().fn["rsgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.0, vp(ce), M)
    rb.lib().reproBLAS_rsgemm(3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.0, vp(cg), M)
    rb._check()
    assert eq_arrays(cg, ce)


def test_gemm_binned_accumulate(gpu):
    """binnedBLAS_sbsgemm accumulates into a float_binned C over two K panels."""
    import reproblas_amd as rb

    M, N, K, fold = 33, 21, 500, 3
    O = oracle_s()
    A = dg.wide_s(M * K, 95)
    B = dg.wide_s(K * N, 96)
    Cb = np.zeros(M * N * 2 * fold, np.float32)
    Ce = Cb.copy()
    for lo, hi in ((0, 200), (200, K)):
        Ap = np.ascontiguousarray(A.reshape(K, M)[lo:hi]).ravel()
        Bp = np.ascontiguousarray(B.reshape(N, K)[:, lo:hi]).ravel()
        rb.sbsgemm(fold, "C", "N", "N", M, N, hi - lo, 1.5, Ap, M, Bp, hi - lo, Cb, M)
        O.fn["sbsgemm"](fold, b"C", b"N", b"N", M, N, hi - lo, 1.5, vp(Ap), M, vp(Bp), hi - lo, vp(Ce), M)
    assert eq_arrays(Cb, Ce)


def test_gemv_binned_accumulate(gpu):
    import reproblas_amd as rb

    M, N, fold = 300, 700, 4
    O = oracle_s()
    A = dg.wide_s(M * N, 97)
    x = dg.wide_s(N, 98)
    for trans in ("N", "T"):
        ny = M if trans == "N" else N
        nx = N if trans == "N" else M
        Yb = np.zeros(ny * 2 * fold, np.float32)
        Ye = Yb.copy()
        for alpha in (1.0, -0.25):
            rb.sbsgemv(fold, "C", trans, M, N, alpha, A, M, x[:nx].copy(), 1, Yb)
            O.fn["sbsgemv"](fold, b"C", trans.encode(), M, N, alpha, vp(A), M, vp(x), 1, vp(Ye), 1)
        assert eq_arrays(Yb, Ye), trans


@pytest.mark.parametrize("fold", [5, 9, 21])
def test_high_fold_gemv_gemm(gpu, fold):
    import reproblas_amd as rb

    O = oracle_s()
    M, N, K = 70, 50, 400
    A = dg.wide_s(M * K, 101)
    B = dg.wide_s(K * N, 102)
    ce, cg = np.zeros(M * N, np.float32), np.zeros(M * N, np.float32)
    O.fn["rsgemm"](fold, b"C", b"N", b"T", M, N, K, 2.0, vp(A), M, vp(B), N, 0.0, vp(ce), M)
    rb.lib().reproBLAS_rsgemm(fold, b"C", b"N", b"T", M, N, K, 2.0, vp(A), M, vp(B), N, 0.0, vp(cg), M)
    rb._check()
    assert eq_arrays(cg, ce)
    x = dg.wide_s(K, 103)
    ye, yg = np.zeros(M, np.float32), np.zeros(M, np.float32)
    O.fn["rsgemv"](fold, b"C", b"N", M, K, 1.0, vp(A), M, vp(x), 1, 0.0, vp(ye), 1)
    rb.lib().reproBLAS_rsgemv(fold, b"C", b"N", M, K, 1.0, vp(A), M, vp(x), 1, 0.0, vp(yg), 1)
    rb._check()
    assert eq_arrays(yg, ye)


@pytest.mark.slow
@pytest.mark.parametrize("fold", [2, 3])
def test_full_size_sdot_snrm2(gpu, fold):
    """N = 2^28 float wide-range dot, sum and nrm2, bit-identical to the oracle."""
    import torch

    import reproblas_amd as rb

    n = 1 << 28
    x = dg.wide_s(n, 201)
    y = dg.wide_s(n, 202)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    O = oracle_s()
    assert same(rb.rsdot(tx, ty, fold), O.rsdot(fold, x, y))
    assert same(rb.rsnrm2(tx, fold), O.rsnrm2(fold, x))
    assert same(rb.rssum(tx, fold), O.rssum(fold, x))


def test_matches_reference_fixtures(gpu):
    """GPU results equal the compiled reference's own outputs stored in
    tests/golden/sc.npz (single precision BLAS1, gemv, gemm)."""
    import os

    import reproblas_amd as rb

    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sc.npz"))
    for k in [k for k in G.files if k.startswith("s_") and k.endswith("_res")]:
        parts = k.split("_")
        kind, n, fold = "_".join(parts[1:-3]), parts[-3], int(parts[-2])
        x, y = G[f"s_{kind}_{n}_x"], G[f"s_{kind}_{n}_y"]
        got = np.array([rb.rssum(x, fold), rb.rsasum(x, fold), rb.rsnrm2(x, fold), rb.rsdot(x, y, fold)], np.float32)
        assert eq_arrays(got, G[k]), k
    A, B, C0 = G["s_mat_A"], G["s_mat_B"], G["s_mat_C"]
    M, N, K = 37, 29, 300
    L = rb.lib()
    for order in "CR":
        for tr in "NT":
            y = C0.copy()
            lda = (M if order == "C" else N) + 1
            L.reproBLAS_rsgemv(3, order.encode(), tr.encode(), M, N, 1.5, vp(A), lda, vp(B), 1, 0.5, vp(y), 1)
            rb._check()
            assert eq_arrays(y, G[f"s_gemv_{order}{tr}"]), (order, tr)
            for tb in "NT":
                rm = order == "R"
                la = (K if tr == "N" else M) if rm else (M if tr == "N" else K)
                lb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
                c = C0.copy()
                L.reproBLAS_rsgemm(3, order.encode(), tr.encode(), tb.encode(), M, N, K, 0.75, vp(A), la, vp(B), lb,
                                   0.5, vp(c), N if rm else M)
                rb._check()
                assert eq_arrays(c, G[f"s_gemm_{order}{tr}{tb}"]), (order, tr, tb)
