"""GPU parity on the edges of the API: calls that form no product, binned
arrays on the device, binned values with separate primary/carry strides, an
empty nrm2 accumulator with a caller scale, concurrent host threads, and the
full-size dgemm checked output by output against one binned dot per element.

Every comparison is bit for bit against the oracle (and, where the CPU build
of the reference is present, against the reference's own strided entry
points). Reference behaviour cited per test.
"""
import concurrent.futures
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle, oracle_s, oracle_z, reference

pytestmark = pytest.mark.gpu


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def u64(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same_bits(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (u64(a) == u64(b))))


def dptr(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + t.element_size() * off)


# -------------------------------------------- gemm / gemv with no products --
# rdgemm.c:50-52 returns for M == 0 or N == 0 without touching C;
# dbdgemm.ccog:77-80 adds nothing for K == 0 or alpha == 0, so C becomes
# ddbconv(dbdconv(beta*C)) (zero when beta == 0) and A, B are never read.
NO_PRODUCT = [("alpha0", 0.0, 0.0, 64), ("alpha0", 0.0, 0.5, 64), ("alpha0", 0.0, 1.0, 64),
              ("alpha0", 0.0, -3.0, 64), ("k0", 1.0, 0.0, 0), ("k0", 2.0, 0.5, 0), ("k0", 1.0, 1.0, 0)]


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("case,alpha,beta,K", NO_PRODUCT)
@pytest.mark.parametrize("fold", [2, 3, 9])
def test_rdgemm_no_products(gpu, order, case, alpha, beta, K, fold):
    import torch

    import reproblas_amd as rb

    M, N = 70, 45
    ldc = (N if order == "R" else M) + 3
    C0 = dg.wide(ldc * (M if order == "R" else N), 501, -200, 200)
    C0[7] = np.nan
    C0[11] = np.inf
    A = dg.normal(max(M, 1) * 64 * 2, 502)
    B = dg.normal(max(N, 1) * 64 * 2, 503)
    lda = M if order == "C" else 64
    ldb = 64 if order == "C" else N
    Ce = C0.copy()
    oracle().fn["rdgemm"](fold, order.encode(), b"N", b"N", M, N, K, alpha, vp(A), max(lda, 1), vp(B), max(ldb, 1),
                          beta, vp(Ce), ldc)
    # host operands; A and B are NULL: the call must not read them
    Cg = C0.copy()
    rb.lib().reproBLAS_rdgemm(fold, order.encode(), b"N", b"N", M, N, K, alpha, None, max(lda, 1), None, max(ldb, 1),
                              beta, vp(Cg), ldc)
    rb._check()
    assert same_bits(Cg, Ce), (case, alpha, beta, K, fold)
    # device API, C in place
    tC = torch.from_numpy(C0.copy()).to(gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert rb.lib().reproblas_hip_rdgemm_device(fold, order.encode(), b"N", b"T", M, N, K, alpha, None, max(lda, 1),
                                                None, max(ldb, 1), beta, dptr(tC), ldc, s) == 0
    assert same_bits(tC.cpu().numpy(), Ce)


@pytest.mark.parametrize("M,N", [(0, 50), (50, 0), (0, 0)])
def test_rdgemm_empty_output(gpu, M, N):
    """rdgemm.c:50-52: M == 0 or N == 0 returns before C is read or written."""
    import reproblas_amd as rb

    C0 = dg.normal(100, 504)
    C0[3] = np.nan
    Cg = C0.copy()
    rb.lib().reproBLAS_rdgemm(3, b"C", b"N", b"N", M, N, 40, 1.0, None, 1, None, 1, 0.5, vp(Cg), 1)
    rb._check()
    assert same_bits(Cg, C0)
    Ce = C0.copy()
    oracle().fn["rdgemm"](3, b"C", b"N", b"N", M, N, 40, 1.0, vp(C0), 1, vp(C0), 1, 0.5, vp(Ce), 1)
    assert same_bits(Ce, C0)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 0.5, 1.0])
def test_rdgemv_alpha0_reads_nothing(gpu, trans, beta):
    """rdgemv.c:48-84 with alpha == 0: y = ddbconv(dbdconv(beta*y)); A, x unread."""
    import reproblas_amd as rb

    M, N = 300, 77
    ly = M if trans == "N" else N
    y0 = dg.wide(ly * 2, 505, -300, 300)
    y0[4] = -np.inf
    A = dg.normal(M * N, 506)
    x = dg.normal(max(M, N), 507)
    for fold in (2, 3, 5):
        ye = y0.copy()
        oracle().fn["rdgemv"](fold, b"C", trans.encode(), M, N, 0.0, vp(A), M, vp(x), 1, beta, vp(ye), 2)
        yg = y0.copy()
        rb.lib().reproBLAS_rdgemv(fold, b"C", trans.encode(), M, N, 0.0, None, M, None, 1, beta, vp(yg), 2)
        rb._check()
        assert same_bits(yg, ye), (trans, beta, fold)


@pytest.mark.parametrize("M,N", [(0, 30), (30, 0)])
def test_rdgemv_empty(gpu, M, N):
    """rdgemv.c:43-45: M == 0 or N == 0 returns before y is touched."""
    import reproblas_amd as rb

    y0 = dg.normal(40, 508)
    yg = y0.copy()
    for trans in (b"N", b"T"):
        rb.lib().reproBLAS_rdgemv(3, b"C", trans, M, N, 1.0, None, 1, None, 1, 0.5, vp(yg), 1)
        rb._check()
        assert same_bits(yg, y0)


def test_float_no_products(gpu):
    """rsgemm / rsgemv with alpha == 0 and K == 0 (the float path of the same kernels)."""
    import reproblas_amd as rb

    O = oracle_s()
    M, N = 33, 21
    C0 = dg.wide(M * N, 509, -50, 50).astype(np.float32)
    for alpha, beta, K in ((0.0, 0.5, 10), (1.0, 2.0, 0), (0.0, 0.0, 10)):
        Ce, Cg = C0.copy(), C0.copy()
        A = dg.normal(M * 10, 510).astype(np.float32)
        O.fn["rsgemm"](3, b"C", b"N", b"N", M, N, K, alpha, vp(A), M, vp(A), 10, beta, vp(Ce), M)
        rb.lib().reproBLAS_rsgemm(3, b"C", b"N", b"N", M, N, K, alpha, None, M, None, 10, beta, vp(Cg), M)
        rb._check()
        assert np.array_equal(Cg.view(np.uint32), Ce.view(np.uint32)), (alpha, beta, K)
        ye, yg = C0[:M].copy(), C0[:M].copy()
        O.fn["rsgemv"](3, b"C", b"N", M, N, 0.0, vp(A), M, vp(A), 1, beta, vp(ye), 1)
        rb.lib().reproBLAS_rsgemv(3, b"C", b"N", M, N, 0.0, None, M, None, 1, beta, vp(yg), 1)
        rb._check()
        assert np.array_equal(yg.view(np.uint32), ye.view(np.uint32))


# ------------------------------------------- binned arrays on the device --
@pytest.mark.parametrize("fold", [2, 3, 4, 9])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 1), (1, 3), (3, 2)])
def test_dbdbaddv_device_arrays(gpu, fold, incx, incy):
    """binned_dbdbaddv (src/binned/dbdbaddv.c:18-23) on device double_binned
    arrays: Y[i*incY] += X[i*incX] elementwise, strides in binned elements."""
    import torch

    import reproblas_amd as rb

    O = oracle()
    n = 1000
    nb = 2 * fold
    X = np.zeros(n * incx * nb)
    Y = np.zeros(n * incy * nb)
    # binned values of mixed magnitude, some at index 0, one NaN and one inf
    vals_x = dg.wide(n, 511, -300, 300)
    vals_y = dg.wide(n, 512, -300, 300) * 1e3
    vals_x[5] = 1.7e308
    vals_y[6] = -1.6e308
    vals_x[9] = np.nan
    vals_y[10] = np.inf
    for i in range(n):
        O.fn["dbdconv"](fold, vals_x[i], ctypes.c_void_p(X.ctypes.data + 8 * i * incx * nb))
        O.fn["dbdconv"](fold, vals_y[i], ctypes.c_void_p(Y.ctypes.data + 8 * i * incy * nb))
        # a second term so the carries and lower bins are populated
        O.fn["dbdadd"](fold, vals_x[(i + 1) % n] * 1e-7, ctypes.c_void_p(X.ctypes.data + 8 * i * incx * nb))
    Ye = Y.copy()
    for i in range(n):
        O.fn["dbdbadd"](fold, ctypes.c_void_p(X.ctypes.data + 8 * i * incx * nb),
                        ctypes.c_void_p(Ye.ctypes.data + 8 * i * incy * nb))
    tX, tY = torch.from_numpy(X).to(gpu), torch.from_numpy(Y.copy()).to(gpu)
    rb.dbdbaddv(tX, tY, n, fold, incx, incy)
    got = tY.cpu().numpy()
    for i in range(n):
        g = got[i * incy * nb:(i * incy + 1) * nb]
        e = Ye[i * incy * nb:(i * incy + 1) * nb]
        if np.isnan(e[0]) or np.isinf(e[0]):
            assert same_bits(g[0], e[0]), i  # only the class is defined for non-finite values
        else:
            assert np.array_equal(u64(g), u64(e)), (i, g, e)
    # untouched gaps between strided elements
    mask = np.ones(len(Y), bool)
    for i in range(n):
        mask[i * incy * nb:(i * incy + 1) * nb] = False
    assert np.array_equal(u64(got[mask]), u64(Y[mask]))
    # the host path gives the same bits
    Yh = Y.copy()
    rb.dbdbaddv(X, Yh, n, fold, incx, incy)
    assert same_bits(Yh, got)


def test_dbdbaddv_mixed_memory_refused(gpu):
    import torch

    import reproblas_amd as rb

    X = np.zeros(6)
    tY = torch.zeros(6, dtype=torch.float64, device=gpu)
    with pytest.raises(rb.ReproBLASError):
        rb.dbdbaddv(X, tY, 1, 3)


# --------------------------------- binned values with their own strides --
@pytest.mark.parametrize("op", ["sum", "asum", "dot", "ssq"])
@pytest.mark.parametrize("incp,incc", [(2, 1), (1, 3), (3, 2), (-1, 1)])
def test_dm_forms_separate_strides(gpu, op, incp, incc):
    """binnedBLAS_dmdsum / dmdasum / dmddot / dmdssq (dmdsum.ccog:57-89 and
    siblings) on primaries and carries with their own strides equal the packed
    dbd* form on the same value (the reference's dbd* calls are exactly the
    dm* calls with strides 1), and the reference's own dm* entry points."""
    import reproblas_amd as rb

    L = rb.lib()
    fold = 3
    n = 5000
    x = dg.wide(n, 520)
    y = dg.wide(n, 521)
    start = dg.wide(3, 522)
    ap, ac = abs(incp), abs(incc)
    for variant in ("gpu", "ref"):
        R = reference() if variant == "ref" else None
        if variant == "ref" and R is None:
            continue
        pri = np.full(fold * ap + 4, 7.25)
        car = np.full(fold * ac + 4, -3.5)
        packed = np.zeros(2 * fold)
        for v in start:
            oracle().fn["dbdadd"](fold, v, vp(packed))
        # lay the packed value out with the strides (negative: not pointer-adjusted,
        # element i at base + i*inc, as the reference indexes it)
        pbase = (fold - 1) * ap if incp < 0 else 0
        cbase = (fold - 1) * ac if incc < 0 else 0
        for i in range(fold):
            pri[pbase + i * incp] = packed[i]
            car[cbase + i * incc] = packed[fold + i]
        pp = ctypes.c_void_p(pri.ctypes.data + 8 * pbase)
        cp = ctypes.c_void_p(car.ctypes.data + 8 * cbase)
        exp = packed.copy()
        scale_e = scale_g = None
        if op == "sum":
            oracle().dbdsum(fold, x, exp)
        elif op == "asum":
            oracle().dbdasum(fold, x, exp)
        elif op == "dot":
            oracle().dbddot(fold, x, y, exp)
        else:
            scale_e = oracle().dbdssq(fold, x, 0.0, exp)
        if variant == "gpu":
            if op == "sum":
                L.binnedBLAS_dmdsum(fold, n, vp(x), 1, pp, incp, cp, incc)
            elif op == "asum":
                L.binnedBLAS_dmdasum(fold, n, vp(x), 1, pp, incp, cp, incc)
            elif op == "dot":
                L.binnedBLAS_dmddot(fold, n, vp(x), 1, vp(y), 1, pp, incp, cp, incc)
            else:
                scale_g = L.binnedBLAS_dmdssq(fold, n, vp(x), 1, ctypes.c_double(0.0), pp, incp, cp, incc)
            rb._check()
        else:
            fn = {"sum": "binnedBLAS_dmdsum", "asum": "binnedBLAS_dmdasum", "dot": "binnedBLAS_dmddot",
                  "ssq": "binnedBLAS_dmdssq"}[op]
            f = getattr(R.lib, fn)
            if op == "ssq":
                f.restype = ctypes.c_double
                scale_g = f(fold, n, vp(x), 1, ctypes.c_double(0.0), pp, incp, cp, incc)
            elif op == "dot":
                f(fold, n, vp(x), 1, vp(y), 1, pp, incp, cp, incc)
            else:
                f(fold, n, vp(x), 1, pp, incp, cp, incc)
        got_p = np.array([pri[pbase + i * incp] for i in range(fold)])
        got_c = np.array([car[cbase + i * incc] for i in range(fold)])
        assert np.array_equal(u64(got_p), u64(exp[:fold])), (variant, op, incp, incc)
        assert np.array_equal(u64(got_c), u64(exp[fold:])), (variant, op, incp, incc)
        if op == "ssq":
            assert same_bits(scale_g, scale_e)
        # slots between the strided elements are untouched
        pm = np.ones(len(pri), bool)
        cm = np.ones(len(car), bool)
        for i in range(fold):
            pm[pbase + i * incp] = False
            cm[cbase + i * incc] = False
        assert np.all(pri[pm] == 7.25) and np.all(car[cm] == -3.5)


# ------------------------- nrm2 into an empty value that carries a scale --
@pytest.mark.parametrize("fold", [2, 3, 4, 7])
@pytest.mark.parametrize("xs", ["small", "zeros", "mixed"])
def test_ssq_empty_accumulator_keeps_caller_scale(gpu, fold, xs):
    """dmdssq.ccog:62-101: with Y == 0 and scaleY > dscale(max|x|) the scale
    stays scaleY and every square is taken at that scale. Host and device Y,
    double and float, and the complex dbzssq / sbcssq."""
    import torch

    import reproblas_amd as rb

    O = oracle()
    if xs == "small":
        x = np.array([1.0, 2.0, 3.0])
    elif xs == "zeros":
        x = np.zeros(5)
    else:
        x = dg.wide(3000, 530)
    for scale_y in (O.dscale(1e60), O.dscale(1e-3), O.dscale(0.0), 0.0):
        Ye = np.zeros(2 * fold)
        se = O.dbdssq(fold, x, scale_y, Ye)
        Yh = np.zeros(2 * fold)
        sh = rb.dbdssq(Yh, x, scale_y, fold)
        assert same_bits(sh, se), (scale_y, sh, se)
        assert np.array_equal(u64(Yh), u64(Ye)), (scale_y, Yh, Ye)
        # device value and device scale
        tY = torch.zeros(2 * fold, dtype=torch.float64, device=gpu)
        tx = torch.from_numpy(x).to(gpu)
        ts = torch.tensor([scale_y], dtype=torch.float64, device=gpu)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert rb.lib().reproblas_hip_dbdssq_device(fold, len(x), dptr(tx), 1, dptr(ts), dptr(tY), s) == 0
        assert same_bits(ts.item(), se)
        assert np.array_equal(u64(tY.cpu().numpy()), u64(Ye))
        # complex: the same values as real parts, zero imaginary parts
        xz = np.zeros(2 * len(x))
        xz[0::2] = x
        OZ = oracle_z("z")
        Yze = np.zeros(2 * fold)
        sze = OZ.bssq(fold, xz, scale_y, Yze)
        Yzg = np.zeros(2 * fold)
        szg = rb.dbzssq(Yzg, xz.view(np.complex128), scale_y, fold)
        assert same_bits(szg, sze)
        assert np.array_equal(u64(Yzg), u64(Yze))
    # float path (float_binned, sbsssq; smsssq.ccog has the same driver)
    OS = oracle_s()
    xf = x.astype(np.float32)
    for scale_y in (float(OS.fn["sscale"](1e30)), float(OS.fn["sscale"](1e-3)), float(OS.fn["sscale"](0.0)), 0.0):
        Yfe = np.zeros(2 * fold, np.float32)
        sfe = OS.sbsssq(fold, xf, scale_y, Yfe)
        Yfg = np.zeros(2 * fold, np.float32)
        sfg = rb.sbsssq(Yfg, xf, scale_y, fold)
        assert np.float32(sfg).view(np.uint32) == np.float32(sfe).view(np.uint32), (scale_y, sfg, sfe)
        assert np.array_equal(Yfg.view(np.uint32), Yfe.view(np.uint32)), (scale_y, Yfg, Yfe)


# ------------------------------------------------ concurrent host threads --
def test_concurrent_host_threads(gpu):
    """Four host threads call reproBLAS_rddot, reproBLAS_rdsum and
    reproBLAS_rdgemm on host operands at once (sharing the null stream, then
    each on its own stream): every result is bit-identical to the oracle."""
    import torch

    import reproblas_amd as rb

    L = rb.lib()
    O = oracle()
    nthreads, reps = 4, 6
    sizes = [(1 << 16) + 17 * t for t in range(nthreads)]
    xs = [dg.wide(n, 540 + t) for t, n in enumerate(sizes)]
    ys = [dg.wide(n, 550 + t) for t, n in enumerate(sizes)]
    exp_dot = [O.rddot(3, x, y) for x, y in zip(xs, ys)]
    exp_sum = [O.rdsum(3, x) for x in xs]
    gm = [(40 + 3 * t, 36 + 5 * t, 300 + 11 * t) for t in range(nthreads)]
    mats = []
    for t, (M, N, K) in enumerate(gm):
        A, B, C = dg.normal(M * K, 560 + t), dg.normal(K * N, 570 + t), dg.normal(M * N, 580 + t)
        Ce = C.copy()
        O.fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.5, vp(Ce), M)
        mats.append((A, B, C, Ce))
    streams = [torch.cuda.Stream() for _ in range(nthreads)]

    def work(t, own_stream):
        if own_stream:
            L.reproblas_hip_set_stream(ctypes.c_void_p(streams[t].cuda_stream))
        else:
            L.reproblas_hip_set_stream(None)
        bad = []
        for r in range(reps):
            d = L.reproBLAS_rddot(3, sizes[t], vp(xs[t]), 1, vp(ys[t]), 1)
            s_ = L.reproBLAS_rdsum(3, sizes[t], vp(xs[t]), 1)
            A, B, C, Ce = mats[t]
            M, N, K = gm[t]
            Cg = C.copy()
            L.reproBLAS_rdgemm(3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.5, vp(Cg), M)
            if L.reproblas_hip_last_error() != 0:
                bad.append(("error", L.reproblas_hip_last_error_string()))
            if not same_bits(d, exp_dot[t]):
                bad.append(("dot", r))
            if not same_bits(s_, exp_sum[t]):
                bad.append(("sum", r))
            if not same_bits(Cg, Ce):
                bad.append(("gemm", r))
        L.reproblas_hip_set_stream(None)
        return bad

    for own in (False, True):
        with concurrent.futures.ThreadPoolExecutor(nthreads) as ex:
            results = list(ex.map(lambda t: work(t, own), range(nthreads)))
        assert all(not r for r in results), (own, results)
    torch.cuda.synchronize()


def test_tiling_is_per_thread(gpu):
    """The gemm tiling set in one host thread does not change another's."""
    import threading

    import reproblas_amd as rb

    L = rb.lib()
    old = L.reproblas_hip_gemm_get_tiling()
    seen = {}

    def other():
        seen["before"] = L.reproblas_hip_gemm_get_tiling()
        L.reproblas_hip_gemm_set_tiling(0)
        seen["after"] = L.reproblas_hip_gemm_get_tiling()

    L.reproblas_hip_gemm_set_tiling(2)
    th = threading.Thread(target=other)
    th.start()
    th.join()
    assert L.reproblas_hip_gemm_get_tiling() == 2
    assert seen["before"] == 1 and seen["after"] == 0  # a new thread starts at the default
    L.reproblas_hip_gemm_set_tiling(old)


# ------------------------------------------- full-size dgemm, per output --
def _binned_dot_per_output(fold, A, lda, ta, B, ldb, tb, K, i, j, beta, c0):
    """C_ij as one binned dot (wrap_rdgemm.h:16-70): dbdconv(beta*c0) (+)
    sum_k fl(A_ik * B_kj) through the oracle's dbddot, then ddbconv."""
    O = oracle()
    Y = np.zeros(2 * fold)
    if beta != 0.0:
        O.fn["dbdconv"](fold, c0 if beta == 1.0 else c0 * beta, vp(Y))
    # op(A) row i and op(B) column j as strided views (column-major storage)
    if ta == "N":
        pa, inca = A.ctypes.data + 8 * i, lda
    else:
        pa, inca = A.ctypes.data + 8 * i * lda, 1
    if tb == "N":
        pb, incb = B.ctypes.data + 8 * j * ldb, 1
    else:
        pb, incb = B.ctypes.data + 8 * j, ldb
    O.fn["dbddot"](fold, K, ctypes.c_void_p(pa), inca, ctypes.c_void_p(pb), incb, vp(Y))
    return O.fn["ddbconv"](fold, vp(Y))


@pytest.mark.slow
def test_full_size_dgemm_per_output(gpu):
    """dgemm 4096^3, fold 3, column-major, NoTrans and Trans A, beta 0 and 0.5:
    1024 sampled outputs per call equal one binned dot per element, and the
    whole 4096^2 result is bit-identical across the three tilings."""
    import torch

    import reproblas_amd as rb

    n = 4096
    A = dg.normal(n * n, 601)
    B = dg.normal(n * n, 602)
    C0 = dg.normal(n * n, 603)
    tA, tB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    g = dg.rng(604)
    L = rb.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    old = rb.get_gemm_tiling()
    try:
        for ta in ("N", "T"):
            for beta in (0.0, 0.5):
                outs = []
                for tiling in (0, 1, 2):
                    rb.set_gemm_tiling(tiling)
                    tC = torch.from_numpy(C0.copy()).to(gpu)
                    assert L.reproblas_hip_rdgemm_device(3, b"C", ta.encode(), b"N", n, n, n, 1.0, dptr(tA), n,
                                                         dptr(tB), n, beta, dptr(tC), n, s) == 0
                    outs.append(tC.cpu().numpy())
                    del tC
                assert np.array_equal(u64(outs[0]), u64(outs[1])) and np.array_equal(u64(outs[0]), u64(outs[2])), \
                    (ta, beta)
                C = outs[0]
                ii = g.integers(0, n, 1024)
                jj = g.integers(0, n, 1024)
                for i, j in zip(ii, jj):
                    e = _binned_dot_per_output(3, A, n, ta, B, n, "N", n, int(i), int(j), beta, C0[j * n + i])
                    assert same_bits(C[j * n + i], e), (ta, beta, int(i), int(j))
    finally:
        rb.set_gemm_tiling(old)


# ------------------------------------------------ tall-skinny dgemm --
@pytest.mark.slow
def test_tall_dgemm_beyond_65535_tile_rows(gpu):
    """rdgemm over 2^22 + 37 rows (more than 65535 row tiles at every tiling,
    past the y-dimension limit of a launch grid), N = 3, K = 5, column-major:
    the whole result equals the oracle's, at each tiling."""
    import torch

    import reproblas_amd as rb

    M, N, K = (1 << 22) + 37, 3, 5
    A = dg.normal(M * K, 611)
    B = dg.normal(K * N, 612)
    C0 = dg.normal(M * N, 613)
    Ce = C0.copy()
    oracle().fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.5, vp(Ce), M)
    tA, tB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    L = rb.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    old = rb.get_gemm_tiling()
    try:
        for tiling in (0, 1, 2):
            rb.set_gemm_tiling(tiling)
            tC = torch.from_numpy(C0).to(gpu)
            assert L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", M, N, K, 1.0, dptr(tA), M, dptr(tB), K, 0.5,
                                                 dptr(tC), M, s) == 0, L.reproblas_hip_last_error_string()
            torch.cuda.synchronize()
            assert np.array_equal(u64(tC.cpu().numpy()), u64(Ce)), tiling
    finally:
        rb.set_gemm_tiling(old)


# ------------------------------------ full-size ragged batch, sampled --
@pytest.mark.slow
def test_full_size_ragged_batch(gpu):
    """The ragged batch at benchmark scale (10,000 dot segments with lengths
    log-uniform in [16, 2^20], 512 gemms with M, N, K in {32, ..., 512}):
    one graph launch gives the same bits as per-call launches into separate
    outputs for every segment and every gemm element, and a sample of each
    equals the oracle."""
    import torch

    import reproblas_amd as rb
    from reproblas_amd.batch import Batch

    L = rb.lib()
    g = dg.rng(7)
    nseg = 10000
    lens = np.exp(g.uniform(np.log(16), np.log(1 << 20), nseg)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(off[-1])
    x = dg.wide(total, 11)
    y = dg.wide(total, 12)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    res = torch.zeros(nseg, dtype=torch.float64, device=gpu)
    res2 = torch.zeros(nseg, dtype=torch.float64, device=gpu)
    sizes = g.integers(1, 17, (512, 3)) * 32
    host = [(dg.normal(int(K) * int(M), 1000 + q), dg.normal(int(N) * int(K), 2000 + q))
            for q, (M, N, K) in enumerate(sizes)]
    triples = [(torch.from_numpy(a).to(gpu).view(int(K), int(M)), torch.from_numpy(b).to(gpu).view(int(N), int(K)),
                torch.zeros(int(N), int(M), dtype=torch.float64, device=gpu))
               for (a, b), (M, N, K) in zip(host, sizes)]
    C2 = [torch.zeros_like(c) for _, _, c in triples]
    b = Batch(3)
    b.add_segments("dot", off, tx, ty, res)
    b.add_gemms("N", "N", triples)
    b.launch(graph=True)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for q in range(nseg):
        assert L.reproblas_hip_rddot_device(3, int(lens[q]), dptr(tx, int(off[q])), 1, dptr(ty, int(off[q])), 1,
                                            dptr(res2, q), s) == 0
    for (M, N, K), (A, B, _), C in zip(sizes, triples, C2):
        assert L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", int(M), int(N), int(K), 1.0, dptr(A), int(M),
                                             dptr(B), int(K), 0.0, dptr(C), int(M), s) == 0
    torch.cuda.synchronize()
    assert torch.equal(res.view(torch.int64), res2.view(torch.int64))
    for (_, _, C), Cc in zip(triples, C2):
        assert torch.equal(C.view(torch.int64), Cc.view(torch.int64))
    O = oracle()
    r = res.cpu().numpy()
    for q in g.choice(nseg, 40, replace=False):
        a, e = int(off[q]), int(off[q + 1])
        assert same_bits(r[q], O.rddot(3, x[a:e], y[a:e])), int(q)
    for q in g.choice(512, 8, replace=False):
        M, N, K = (int(v) for v in sizes[q])
        Ce = np.zeros(M * N)
        a, bb = host[q]
        O.fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(a), M, vp(bb), K, 0.0, vp(Ce), M)
        assert same_bits(triples[q][2].cpu().numpy().ravel(), Ce), int(q)
    b.close()
