"""GPU parity of gemv and gemm against the oracle (bit-identical outputs).

Reference behaviour: src/reproBLAS/rdgemv.c:35-87 with dbdgemv.ccog
(alpha on x) and src/reproBLAS/rdgemm.c:41-105 with dbdgemm.ccog (alpha on
A); beta == 0 ignores the output, beta == 1 converts it unscaled.
"""
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle

pytestmark = pytest.mark.gpu


def eq_arrays(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint64) == b.view(np.uint64))))


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (255, 512), (300, 77), (1000, 2049)])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (2.0, 0.5), (0.5, 1.0), (0.0, 2.0)])
def test_rdgemv(gpu, order, trans, shape, ab):
    import reproblas_amd as rb

    M, N = shape
    alpha, beta = ab
    lda = (M if order == "C" else N) + 3
    A = dg.normal(lda * (N if order == "C" else M), 61)
    lx = N if trans == "N" else M
    ly = M if trans == "N" else N
    x = dg.normal(lx, 62)
    y0 = dg.normal(ly, 63)
    for fold in (2, 3, 4):
        ye = y0.copy()
        oracle().fn["rdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta, vp(ye), 1)
        yg = y0.copy()
        rb.lib().reproBLAS_rdgemv(fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                  vp(yg), 1)
        rb._check()
        assert eq_arrays(yg, ye), (order, trans, shape, ab, fold)


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (64, 64, 64), (96, 80, 72), (33, 65, 130), (128, 40, 2100)])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (2.0, 0.5), (0.5, 1.0)])
def test_rdgemm(gpu, order, ta, tb, mnk, ab):
    import reproblas_amd as rb

    M, N, K = mnk
    alpha, beta = ab
    rm = order == "R"
    ar, ac = (M, K) if ta == "N" else (K, M)
    br, bc = (K, N) if tb == "N" else (N, K)
    lda = (ac if rm else ar) + 1
    ldb = (bc if rm else br) + 2
    ldc = (N if rm else M) + 1
    A = dg.normal(lda * (ar if rm else ac), 71)
    B = dg.normal(ldb * (br if rm else bc), 72)
    C0 = dg.normal(ldc * (M if rm else N), 73)
    for fold in (3,) if M * N * K > 100000 else (2, 3, 4):
        Ce = C0.copy()
        oracle().fn["rdgemm"](fold, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B), ldb,
                              beta, vp(Ce), ldc)
        Cg = C0.copy()
        rb.lib().reproBLAS_rdgemm(fold, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B),
                                  ldb, beta, vp(Cg), ldc)
        rb._check()
        assert eq_arrays(Cg, Ce), (order, ta, tb, mnk, ab, fold)


@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_gemm_tilings_identical(gpu, tiling):
    """Every tiling gives the same bits (and the oracle's)."""
    import torch

    import reproblas_amd as rb

    M, N, K = 160, 96, 700
    A = dg.wide(M * K, 81).reshape(K, M)  # column-major M x K
    B = dg.wide(K * N, 82).reshape(N, K)
    C0 = dg.normal((N, M), 83)
    Ce = C0.copy()
    oracle().fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.5, vp(Ce), M)
    old = rb.get_gemm_tiling()
    try:
        rb.set_gemm_tiling(tiling)
        tA, tB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
        tC = torch.from_numpy(C0.copy()).to(gpu)
        rb.rdgemm("N", "N", 1.0, tA, tB, 0.5, tC)
        assert eq_arrays(tC.cpu().numpy(), Ce)
    finally:
        rb.set_gemm_tiling(old)


def test_gemm_special_values(gpu):
    import reproblas_amd as rb

    M, N, K = 40, 36, 50
    A = dg.normal((K, M), 91)
    B = dg.normal((N, K), 92)
    A[3, 5] = np.inf
    A[7, 9] = -np.inf
    B[2, 11] = np.nan
    A[10, 20] = 1e300
    B[4, 10] = 1e300  # overflowing product
    C0 = dg.normal((N, M), 93)
    Ce = C0.copy()
    oracle().fn["rdgemm"](3, b"C", b"N", b"N", M, N, K, 1.0, vp(A), M, vp(B), K, 0.0, vp(Ce), M)
    Cg = C0.copy()
    rb.rdgemm("N", "N", 1.0, A, B, 0.0, Cg)
    assert eq_arrays(Cg, Ce)


def test_gemm_binned_accumulate(gpu):
    """binnedBLAS_dbdgemm accumulates into a binned C; K split in two calls
    equals one call (corroborate_rdgemm.c-style blocking)."""
    import reproblas_amd as rb

    M, N, K = 37, 29, 300
    fold = 3
    A = dg.wide(M * K, 95)
    B = dg.wide(K * N, 96)
    Cb_e = np.zeros(M * N * 2 * fold)
    oracle().dbdgemm(fold, "C", "N", "N", M, N, K, 1.0, A, M, B, K, Cb_e, M)
    Cb = np.zeros(M * N * 2 * fold)
    k1 = 123
    rb.dbdgemm(fold, "C", "N", "N", M, N, k1, 1.0, A, M, B, K, Cb, M)
    rb.dbdgemm(fold, "C", "N", "N", M, N, K - k1, 1.0, A[M * k1:], M, B[k1:], K, Cb, M)
    assert np.array_equal(Cb.view(np.uint64), Cb_e.view(np.uint64))


@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("fold", [3, 4])
def test_gemv_binned_accumulate(gpu, order, fold):
    """binnedBLAS_dbdgemv accumulates into binned outputs (dbdgemv.ccog:57-162):
    two calls with different alpha equal the oracle's two calls, for both
    memory patterns (terms contiguous and lda apart) and an odd output count."""
    import reproblas_amd as rb

    M, N = 301, 700
    lda = N if order == "R" else M
    O = oracle()
    A = dg.wide(M * N, 97)
    x = dg.wide(max(M, N), 98)
    for trans in ("N", "T"):
        ny = M if trans == "N" else N
        nx = N if trans == "N" else M
        Yb = np.zeros(ny * 2 * fold)
        Ye = Yb.copy()
        for alpha in (1.0, -0.25):
            rb.dbdgemv(fold, order, trans, M, N, alpha, A, lda, x[:nx].copy(), 1, Yb)
            O.fn["dbdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, vp(Ye), 1)
        assert eq_arrays(Yb, Ye), (order, trans)


@pytest.mark.slow
def test_gemv_full_size(gpu):
    """Benchmark scale: 16384^2 column-major N and T, fold 3, beta in {0, 0.5}."""
    import torch

    import reproblas_amd as rb

    n = 16384
    A = dg.normal((n, n), 201)  # column-major n x n (transposed view of C order)
    x = dg.normal(n, 202)
    y0 = dg.normal(n, 203)
    tA, tx = torch.from_numpy(A).to(gpu), torch.from_numpy(x).to(gpu)
    O = oracle()
    for trans in ("N", "T"):
        for beta in (0.0, 0.5):
            ty = torch.from_numpy(y0.copy()).to(gpu)
            rb.rdgemv(trans, 1.0, tA, tx, beta, ty)
            ye = y0.copy()
            O.fn["rdgemv"](3, b"C", trans.encode(), n, n, 1.0, vp(A), n, vp(x), 1, beta, vp(ye), 1)
            assert eq_arrays(ty.cpu().numpy(), ye), (trans, beta)


@pytest.mark.parametrize("fold", [5, 9, 20, 52])
def test_high_fold_gemv_gemm(gpu, fold):
    import reproblas_amd as rb

    O = oracle()
    M, N, K = 45, 37, 150
    A = dg.wide(M * K, 1)
    B = dg.wide(K * N, 2)
    C0 = dg.normal(M * N, 3)
    Ce, Cg = C0.copy(), C0.copy()
    O.fn["rdgemm"](fold, b"C", b"N", b"T", M, N, K, 1.0, vp(A), M, vp(B), N, 0.5, vp(Ce), M)
    rb.lib().reproBLAS_rdgemm(fold, b"C", b"N", b"T", M, N, K, 1.0, vp(A), M, vp(B), N, 0.5, vp(Cg), M)
    rb._check()
    assert eq_arrays(Cg, Ce)
    x = dg.normal(K, 4)
    for trans, lx, ly in (("N", K, M), ("T", M, K)):
        y0 = dg.normal(ly, 5)
        ye, yg = y0.copy(), y0.copy()
        xv = x if lx == K else dg.normal(M, 6)
        O.fn["rdgemv"](fold, b"C", trans.encode(), M, K, 2.0, vp(A), M, vp(xv), 1, 0.5, vp(ye), 1)
        rb.lib().reproBLAS_rdgemv(fold, b"C", trans.encode(), M, K, 2.0, vp(A), M, vp(xv), 1, 0.5, vp(yg), 1)
        rb._check()
        assert eq_arrays(yg, ye), trans


@pytest.mark.slow
def test_gemv_beyond_int_elements(gpu):
    """46341 x 46341 column-major A (2^31 + 88 elements, 17 GB: element
    offsets past the int range while M, N and lda fit it), A and x all ones:
    every output of N (col-axpy) and T (row-dot) is exactly 46341, and
    beta = 0.5 adds 0.5 * y exactly."""
    import torch

    import reproblas_amd as rb

    m = 46341
    A = torch.ones((m, m), device=gpu, dtype=torch.float64)
    x = torch.ones(m, device=gpu, dtype=torch.float64)
    for trans in ("N", "T"):
        y = torch.full((m,), 2.0, device=gpu, dtype=torch.float64)
        rb.rdgemv(trans, 1.0, A, x, 0.5, y)
        assert torch.equal(y, torch.full((m,), float(m) + 1.0, device=gpu, dtype=torch.float64)), trans
    del A
    torch.cuda.empty_cache()


@pytest.mark.slow
def test_gemm_beyond_int_elements(gpu):
    """rdgemm with a 46341 x 46341 column-major C (2^31 + 88 elements, 17 GB),
    K = 8, A and B all ones, beta = 0.5 on C = 2: every output is exactly
    8 + 1 = 9 (output offsets past the int range)."""
    import torch

    import reproblas_amd as rb

    m, k = 46341, 8
    A = torch.ones((k, m), device=gpu, dtype=torch.float64)  # m x k column-major, lda = m
    B = torch.ones((m, k), device=gpu, dtype=torch.float64)  # k x m column-major, ldb = k
    C = torch.full((m, m), 2.0, device=gpu, dtype=torch.float64)
    rb.rdgemm("N", "N", 1.0, A, B, 0.5, C, M=m, N=m, K=k, lda=m, ldb=k, ldc=m)
    assert bool((C == 9.0).all())
    del C
    torch.cuda.empty_cache()
