"""GPU parity of the complex path (double complex "z" and float complex "c")
against the complex oracle: bit-identical components (NaN by class).

Reference behaviour: src/reproBLAS/rz*_sub.c, rdzasum.c, rdznrm2.c and the c
forms; src/binnedBLAS/zmz*.ccog, dmz*.ccog, zbzgemv.ccog, zbzgemm.ccog;
src/reproBLAS/rzgemv.c, rzgemm.c (see the oracle's pinning tests for how the
oracle is tied to the compiled reference).
"""
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle_z

pytestmark = pytest.mark.gpu

PRECS = ["z", "c"]


def cdt(prec):
    return np.complex128 if prec == "z" else np.complex64


def rdt(prec):
    return np.float64 if prec == "z" else np.float32


def bits_eq(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    if np.iscomplexobj(a):
        a = a.view(a.real.dtype)
    if np.iscomplexobj(b):
        b = b.view(b.real.dtype)
    assert a.dtype == b.dtype, (a.dtype, b.dtype)
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(u) == b.view(u))))


def cvec(prec, kind, n, seed):
    """n complex values of the precision (numpy complex array)."""
    dt = rdt(prec)
    if kind == "wide":
        v = dg.wide(2 * n, seed, -20, 20)
    elif kind == "uniform":
        v = dg.uniform(2 * n, seed)
    elif kind == "ill":
        v = dg.ill_conditioned(2 * n, seed)
    elif kind == "real_only":
        v = dg.wide(2 * n, seed, -20, 20)
        v[1::2] = 0
    elif prec == "z":
        v = dg.special(kind, 2 * n, seed)
    else:
        v = dg.special_s(kind, 2 * n, seed)
    return np.ascontiguousarray(np.asarray(v, dtype=dt)).view(cdt(prec))


def iv(a):
    """interleaved real view for the oracle"""
    return a.view(a.real.dtype)


def gpu_op(rb, prec, op, fold, x, y=None):
    if prec == "z":
        f = {"sum": rb.rzsum, "dotu": rb.rzdotu, "dotc": rb.rzdotc, "asum": rb.rdzasum, "nrm2": rb.rdznrm2}[op]
    else:
        f = {"sum": rb.rcsum, "dotu": rb.rcdotu, "dotc": rb.rcdotc, "asum": rb.rscasum, "nrm2": rb.rscnrm2}[op]
    r = f(x, y, fold) if op.startswith("dot") else f(x, fold)
    if op in ("asum", "nrm2"):
        return np.array(r, rdt(prec))
    return np.array([r], cdt(prec))


def orc_op(prec, op, fold, x, y=None):
    O = oracle_z(prec)
    if op == "sum":
        return O.rsum(fold, iv(x))
    if op in ("dotu", "dotc"):
        return O.rdot(op[-1], fold, iv(x), iv(y))
    if op == "asum":
        return np.array(O.rasum(fold, iv(x)), rdt(prec))
    return np.array(O.rnrm2(fold, iv(x)), rdt(prec))


OPS = ["sum", "dotu", "dotc", "asum", "nrm2"]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("fold", [2, 3, 4, 8])
@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 4095, 65537, 1 << 20])
def test_random_host_operands(gpu, prec, op, fold, n):
    import reproblas_amd as rb

    x, y = cvec(prec, "wide", n, 11 + n), cvec(prec, "wide", n, 12 + n)
    assert bits_eq(gpu_op(rb, prec, op, fold, x, y), orc_op(prec, op, fold, x, y)), (prec, op, fold, n)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("dist", ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "real_only"])
def test_distributions_device_operands(gpu, prec, op, dist):
    import torch

    import reproblas_amd as rb

    n = 100003
    x, y = cvec(prec, dist, n, 1), cvec(prec, "uniform" if dist != "zeros" else "zeros", n, 2)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    for fold in (2, 3, 4):
        assert bits_eq(gpu_op(rb, prec, op, fold, tx, ty), orc_op(prec, op, fold, x, y)), (prec, op, dist, fold)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("kind", ["nan", "pinf", "ninf", "both_inf"])
@pytest.mark.parametrize("n", [1, 300, 70001])
def test_special_values(gpu, prec, op, kind, n):
    import reproblas_amd as rb

    x, y = cvec(prec, kind, n, 5), cvec(prec, "uniform", n, 6)
    assert bits_eq(gpu_op(rb, prec, op, 3, x, y), orc_op(prec, op, 3, x, y)), (prec, op, kind, n)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("inc", [2, 3, -1, -2])
def test_strides(gpu, prec, inc):
    import reproblas_amd as rb

    n = 20001
    base = cvec(prec, "wide", n * abs(inc), 7)
    O = oracle_z(prec)
    # negative increments: the pointer is the lowest-addressed element
    off = 0 if inc > 0 else (n - 1) * abs(inc) * base.itemsize
    ptr = ctypes.c_void_p(base.ctypes.data + off)
    L = rb.lib()
    rb._bind_stream()
    z = prec
    rd = "rdz" if z == "z" else "rsc"
    for k in ("asum", "nrm2"):
        e = O.fn["r" + k](3, n, ptr, inc)
        r = getattr(L, f"reproBLAS_{rd}{k}")(3, n, ptr, inc)
        rb._check()
        assert bits_eq(np.array(r, rdt(z)), np.array(e, rdt(z))), (k, inc)
    out_g, out_e = np.zeros(2, rdt(z)), np.zeros(2, rdt(z))
    O.fn["rsum"](3, n, ptr, inc, O.p(out_e))
    getattr(L, f"reproBLAS_r{z}sum_sub")(3, n, ptr, inc, ctypes.c_void_p(out_g.ctypes.data))
    rb._check()
    assert bits_eq(out_g, out_e)
    for k in ("u", "c"):
        O.fn["rdot" + k](3, n, ptr, inc, ptr, inc, O.p(out_e))
        getattr(L, f"reproBLAS_r{z}dot{k}_sub")(3, n, ptr, inc, ptr, inc, ctypes.c_void_p(out_g.ctypes.data))
        rb._check()
        assert bits_eq(out_g, out_e), k


@pytest.mark.parametrize("prec", PRECS)
def test_accumulate_into_binned(gpu, prec):
    """zbz*/cbc* accumulate into complex binned values and dbzasum/dbzssq into
    real ones, host and device states, over three chunks."""
    import torch

    import reproblas_amd as rb

    O = oracle_z(prec)
    dt = rdt(prec)
    chunks = [cvec(prec, "wide", 5000, 21), cvec(prec, "wide", 7000, 22) * dt(1e6), cvec(prec, "real_only", 900, 23)]
    ys = [cvec(prec, "wide", len(c), 30 + k) for k, c in enumerate(chunks)]
    z = prec
    for fold in (2, 3, 4):
        for kind in ("sum", "u", "c"):
            Yh, Ye = np.zeros(4 * fold, dt), np.zeros(4 * fold, dt)
            Yd = torch.zeros(4 * fold, dtype=torch.float64 if z == "z" else torch.float32, device=gpu)
            for x, y in zip(chunks, ys):
                fn = getattr(rb, f"{z}b{z}sum" if kind == "sum" else f"{z}b{z}dot{kind}")
                if kind == "sum":
                    fn(Yh, x, fold)
                    fn(Yd, torch.from_numpy(x).to(gpu), fold)
                    O.bsum(fold, iv(x), Ye)
                else:
                    fn(Yh, x, y, fold)
                    fn(Yd, torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu), fold)
                    O.bdot(kind, fold, iv(x), iv(y), Ye)
            assert bits_eq(Yh, Ye), (fold, kind)
            assert bits_eq(Yd.cpu().numpy(), Ye), (fold, kind)
        Ah, Ae = np.zeros(2 * fold, dt), np.zeros(2 * fold, dt)
        Qh, Qe = np.zeros(2 * fold, dt), np.zeros(2 * fold, dt)
        sh = se = 0.0
        for x in chunks:
            getattr(rb, "dbzasum" if z == "z" else "sbcasum")(Ah, x, fold)
            O.basum(fold, iv(x), Ae)
            sh = getattr(rb, "dbzssq" if z == "z" else "sbcssq")(Qh, x, sh, fold)
            se = O.bssq(fold, iv(x), se, Qe)
            assert bits_eq(Ah, Ae) and bits_eq(Qh, Qe) and bits_eq(np.array(sh, dt), np.array(se, dt)), fold


@pytest.mark.parametrize("prec", PRECS)
def test_device_result_api(gpu, prec):
    import torch

    import reproblas_amd as rb

    n = 1 << 17
    x, y = cvec(prec, "wide", n, 41), cvec(prec, "wide", n, 42)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    tdt = torch.float64 if prec == "z" else torch.float32
    out = torch.zeros(8, dtype=tdt, device=gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = rb.lib()
    es = out.element_size()
    p = lambda t, off=0: ctypes.c_void_p(t.data_ptr() + es * off)
    z = prec
    rd = "dz" if z == "z" else "sc"
    assert getattr(L, f"reproblas_hip_r{z}sum_device")(3, n, p(tx), 1, p(out, 0), s) == 0
    assert getattr(L, f"reproblas_hip_r{z}dotu_device")(3, n, p(tx), 1, p(ty), 1, p(out, 2), s) == 0
    assert getattr(L, f"reproblas_hip_r{z}dotc_device")(3, n, p(tx), 1, p(ty), 1, p(out, 4), s) == 0
    assert getattr(L, f"reproblas_hip_r{rd}asum_device")(3, n, p(tx), 1, p(out, 6), s) == 0
    assert getattr(L, f"reproblas_hip_r{rd}nrm2_device")(3, n, p(tx), 1, p(out, 7), s) == 0
    got = out.cpu().numpy()
    assert bits_eq(got[0:2], orc_op(z, "sum", 3, x))
    assert bits_eq(got[2:4], orc_op(z, "dotu", 3, x, y))
    assert bits_eq(got[4:6], orc_op(z, "dotc", 3, x, y))
    assert bits_eq(got[6], orc_op(z, "asum", 3, x))
    assert bits_eq(got[7], orc_op(z, "nrm2", 3, x))


@pytest.mark.parametrize("prec", PRECS)
def test_binned_primitives_and_vector_add(gpu, prec):
    """binned_zbzbaddv on device arrays against zbzbadd per element in the
    oracle; host complex primitives (conv, add, conversion) against it too."""
    import torch

    import reproblas_amd as rb

    O = oracle_z(prec)
    dt = rdt(prec)
    z = prec
    L = rb.lib()
    n, fold = 700, 3
    X = np.zeros((n, 4 * fold), dt)
    Y = np.zeros((n, 4 * fold), dt)
    v = cvec(prec, "wide", 3 * n, 61)
    for i in range(n):
        X[i] = O.bconv(fold, v[i])
        Y[i] = O.bconv(fold, v[n + i])
    fn = getattr(L, f"binned_{z}b{z}baddv")
    fn.restype = None
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    tX, tY = torch.from_numpy(X).to(gpu), torch.from_numpy(Y).to(gpu)
    rb._bind_stream(tX)
    fn(fold, n, ctypes.c_void_p(tX.data_ptr()), 1, ctypes.c_void_p(tY.data_ptr()), 1)
    torch.cuda.synchronize()
    for i in range(n):
        O.fn["bbadd"](fold, O.p(X[i]), O.p(Y[i]))
    assert bits_eq(tY.cpu().numpy(), Y)
    conv = getattr(L, f"binned_{z}b{z}conv")
    conv.restype = None
    conv.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    back = getattr(L, f"binned_{z}{z}bconv_sub")
    back.restype = None
    back.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    for c in v[:40]:
        sv = np.array([c.real, c.imag], dt)
        B = np.zeros(4 * fold, dt)
        conv(fold, ctypes.c_void_p(sv.ctypes.data), ctypes.c_void_p(B.ctypes.data))
        assert bits_eq(B, O.bconv(fold, c))
        out = np.zeros(2, dt)
        back(fold, ctypes.c_void_p(B.ctypes.data), ctypes.c_void_p(out.ctypes.data))
        assert bits_eq(out, O.zbconv(fold, B))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("fold", [9, 13])
def test_high_folds(gpu, prec, fold):
    import reproblas_amd as rb

    for n in (1, 777, 70001):
        x, y = cvec(prec, "wide", n, 300 + n), cvec(prec, "wide", n, 400 + n)
        for op in OPS:
            assert bits_eq(gpu_op(rb, prec, op, fold, x, y), orc_op(prec, op, fold, x, y)), (fold, n, op)


# ---------------------------------------------------------------- gemv/gemm --
ALPHAS = [complex(1, 0), complex(2, -0.5), complex(0, 1)]
BETAS = [complex(0, 0), complex(1, 0), complex(0.5, 0.25)]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (131, 97), (300, 1030)])
def test_rzgemv(gpu, prec, order, trans, shape):
    import reproblas_amd as rb

    O = oracle_z(prec)
    M, N = shape
    lda = (M if order == "C" else N) + 3
    A = cvec(prec, "wide", lda * max(M, N), 61)
    lx = N if trans == "N" else M
    ly = M if trans == "N" else N
    x = cvec(prec, "uniform", lx, 62)
    y0 = cvec(prec, "uniform", ly, 63)
    L = rb.lib()
    for fold in (2, 3):
        for alpha in ALPHAS:
            for beta in BETAS:
                al, be = O.scalar(alpha), O.scalar(beta)
                ye, yg = y0.copy(), y0.copy()
                O.fn["rgemv"](fold, order.encode(), trans.encode(), M, N, O.p(al), O.p(iv(A)), lda, O.p(iv(x)), 1,
                              O.p(be), O.p(iv(ye)), 1)
                getattr(L, f"reproBLAS_r{prec}gemv")(fold, order.encode(), trans.encode(), M, N, O.p(al),
                                                     O.p(iv(A)), lda, O.p(iv(x)), 1, O.p(be), O.p(iv(yg)), 1)
                rb._check()
                assert bits_eq(yg, ye), (order, trans, shape, fold, alpha, beta)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("order", ["C", "R"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (33, 20, 130), (64, 48, 300)])
def test_rzgemm(gpu, prec, order, ta, tb, mnk):
    import reproblas_amd as rb

    O = oracle_z(prec)
    M, N, K = mnk
    rm = order == "R"
    lda = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
    ldb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
    ldc = N if rm else M
    A = cvec(prec, "wide", M * K + 8, 71)
    B = cvec(prec, "wide", K * N + 8, 72)
    C0 = cvec(prec, "uniform", M * N + 8, 73)
    L = rb.lib()
    for alpha, beta in ((complex(1, 0), complex(0, 0)), (complex(1.5, -2), complex(0.5, 0.5)),
                        (complex(1, 0), complex(1, 0)), (complex(0, 0), complex(2, 0))):
        al, be = O.scalar(alpha), O.scalar(beta)
        ce, cg = C0.copy(), C0.copy()
        O.fn["rgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, O.p(al), O.p(iv(A)), lda, O.p(iv(B)), ldb,
                      O.p(be), O.p(iv(ce)), ldc)
        getattr(L, f"reproBLAS_r{prec}gemm")(3, order.encode(), ta.encode(), tb.encode(), M, N, K, O.p(al),
                                             O.p(iv(A)), lda, O.p(iv(B)), ldb, O.p(be), O.p(iv(cg)), ldc)
        rb._check()
        assert bits_eq(cg, ce), (order, ta, tb, mnk, alpha, beta)


@pytest.mark.parametrize("prec", PRECS)
def test_binned_gemm_gemv_accumulate(gpu, prec):
    """zbzgemm / zbzgemv accumulate into complex binned outputs over two calls."""
    import reproblas_amd as rb

    O = oracle_z(prec)
    dt = rdt(prec)
    L = rb.lib()
    z = prec
    M, N, K, fold = 30, 21, 250, 3
    A = cvec(prec, "wide", M * K, 81)
    B = cvec(prec, "wide", K * N, 82)
    Cb = np.zeros(M * N * 4 * fold, dt)
    Ce = Cb.copy()
    for alpha in (complex(1, 0), complex(0.5, -1.5)):
        al = O.scalar(alpha)
        getattr(L, f"binnedBLAS_{z}b{z}gemm")(fold, b"C", b"C", b"N", M, N, K, O.p(al), O.p(iv(A)), K, O.p(iv(B)),
                                              K, O.p(Cb), M)
        rb._check()
        O.fn["bgemm"](fold, b"C", b"C", b"N", M, N, K, O.p(al), O.p(iv(A)), K, O.p(iv(B)), K, O.p(Ce), M)
    assert bits_eq(Cb, Ce)
    x = cvec(prec, "wide", K, 83)
    Yb = np.zeros(M * 4 * fold, dt)
    Ye = Yb.copy()
    for alpha in (complex(1, 0), complex(-2, 0.25)):
        al = O.scalar(alpha)
        getattr(L, f"binnedBLAS_{z}b{z}gemv")(fold, b"R", b"N", M, K, O.p(al), O.p(iv(A)), K, O.p(iv(x)), 1,
                                              O.p(Yb), 1)
        rb._check()
        O.fn["bgemv"](fold, b"R", b"N", M, K, O.p(al), O.p(iv(A)), K, O.p(iv(x)), 1, O.p(Ye), 1)
    assert bits_eq(Yb, Ye)


@pytest.mark.parametrize("prec", PRECS)
def test_device_gemm_gemv(gpu, prec):
    """reproblas_hip_r{z,c}gemm_device / gemv_device on torch tensors."""
    import torch

    import reproblas_amd as rb

    O = oracle_z(prec)
    M, N, K = 96, 80, 200
    A = cvec(prec, "wide", M * K, 91).reshape(K, M)
    B = cvec(prec, "wide", K * N, 92).reshape(N, K)
    C = cvec(prec, "uniform", M * N, 93).reshape(N, M)
    ce = C.copy()
    al, be = O.scalar(complex(0.75, 0.5)), O.scalar(complex(0.5, -1))
    O.fn["rgemm"](3, b"C", b"N", b"T", M, N, K, O.p(al), O.p(iv(A.ravel())), M, O.p(iv(B.ravel())), N, O.p(be),
                  O.p(iv(ce.ravel())), M)
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(a)).to(gpu) for a in (A, B, C))
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = rb.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert getattr(L, f"reproblas_hip_r{prec}gemm_device")(3, b"C", b"N", b"T", M, N, K, O.p(al), vp(tA), M, vp(tB),
                                                           N, O.p(be), vp(tC), M, s) == 0
    assert bits_eq(tC.cpu().numpy(), ce)
    x = cvec(prec, "wide", K, 94)
    y = cvec(prec, "uniform", M, 95)
    ye = y.copy()
    O.fn["rgemv"](3, b"C", b"N", M, K, O.p(al), O.p(iv(A.ravel())), M, O.p(iv(x)), 1, O.p(be), O.p(iv(ye)), 1)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    assert getattr(L, f"reproblas_hip_r{prec}gemv_device")(3, b"C", b"N", M, K, O.p(al), vp(tA), M, vp(tx), 1,
                                                           O.p(be), vp(ty), 1, s) == 0
    assert bits_eq(ty.cpu().numpy(), ye)


@pytest.mark.parametrize("prec", PRECS)
def test_gemv_columns_unaligned_matrix(gpu, prec):
    """Column-layout complex gemv with A at an address aligned only to its
    real type (element loads instead of vector loads): same bits."""
    import torch

    import reproblas_amd as rb

    O = oracle_z(prec)
    M, N = 257, 300
    A = cvec(prec, "wide", M * N, 96)
    x = cvec(prec, "wide", N, 97)
    y = cvec(prec, "uniform", M, 98)
    ye = y.copy()
    al, be = O.scalar(complex(1.25, -0.5)), O.scalar(complex(0.5, 0))
    O.fn["rgemv"](3, b"C", b"N", M, N, O.p(al), O.p(iv(A)), M, O.p(iv(x)), 1, O.p(be), O.p(iv(ye)), 1)
    tdt = torch.float64 if prec == "z" else torch.float32
    buf = torch.zeros(2 * M * N + 1, dtype=tdt, device=gpu)
    buf[1:] = torch.from_numpy(iv(A).copy()).to(gpu)
    tx = torch.from_numpy(iv(x).copy()).to(gpu)
    ty = torch.from_numpy(iv(y).copy()).to(gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = rb.lib()
    pa = ctypes.c_void_p(buf.data_ptr() + buf.element_size())
    assert getattr(L, f"reproblas_hip_r{prec}gemv_device")(3, b"C", b"N", M, N, O.p(al), pa, M,
                                                           ctypes.c_void_p(tx.data_ptr()), 1, O.p(be),
                                                           ctypes.c_void_p(ty.data_ptr()), 1, s) == 0
    assert bits_eq(ty.cpu().numpy(), iv(ye))


@pytest.mark.parametrize("prec", PRECS)
def test_matches_reference_fixtures(gpu, prec):
    """GPU results equal the compiled reference's own complex outputs stored in
    tests/golden/sc.npz (BLAS1 over eight fills, gemv, all gemm layouts)."""
    import os

    import reproblas_amd as rb

    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sc.npz"))
    for k in [k for k in G.files if k.startswith(prec + "_") and k.endswith("_res")]:
        parts = k.split("_")
        kind, n, fold = "_".join(parts[1:-3]), parts[-3], int(parts[-2])
        x, y = G[f"{prec}_{kind}_{n}_x"].view(cdt(prec)), G[f"{prec}_{kind}_{n}_y"].view(cdt(prec))
        got = np.concatenate([iv(gpu_op(rb, prec, "sum", fold, x)), iv(gpu_op(rb, prec, "dotu", fold, x, y)),
                              iv(gpu_op(rb, prec, "dotc", fold, x, y)),
                              np.array([gpu_op(rb, prec, "asum", fold, x), gpu_op(rb, prec, "nrm2", fold, x)])])
        assert bits_eq(got, G[k]), k
    A, B, C = G[f"{prec}_mat_A"], G[f"{prec}_mat_B"], G[f"{prec}_mat_C"]
    dt = A.dtype
    M, N, K = 23, 17, 200
    al, be = np.array([1.5, -0.5], dt), np.array([0.5, 0.25], dt)
    L = rb.lib()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    for order in "CR":
        for ta in "NTC":
            if not (order == "C" and ta == "C"):
                y = C.copy()
                lda = (M if order == "C" else N) + 1
                getattr(L, f"reproBLAS_r{prec}gemv")(3, order.encode(), ta.encode(), M, N, P(al), P(A), lda, P(B), 1,
                                                     P(be), P(y), 1)
                rb._check()
                assert bits_eq(y, G[f"{prec}_gemv_{order}{ta}"]), (order, ta)
            for tb in "NTC":
                rm = order == "R"
                la = (K if ta == "N" else M) if rm else (M if ta == "N" else K)
                lb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
                c = C.copy()
                getattr(L, f"reproBLAS_r{prec}gemm")(3, order.encode(), ta.encode(), tb.encode(), M, N, K, P(al), P(A),
                                                     la, P(B), lb, P(be), P(c), N if rm else M)
                rb._check()
                assert bits_eq(c, G[f"{prec}_gemm_{order}{ta}{tb}"]), (order, ta, tb)


@pytest.mark.slow
def test_full_size_zdotc_zsum_dznrm2(gpu):
    """Benchmark scale: N = 2^27 double complex wide-range zdotc, zsum and
    dznrm2, bit-identical to the oracle."""
    import torch

    import reproblas_amd as rb

    n = 1 << 27
    x, y = cvec("z", "wide", n, 501), cvec("z", "wide", n, 502)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    assert bits_eq(gpu_op(rb, "z", "dotc", 3, tx, ty), orc_op("z", "dotc", 3, x, y))
    assert bits_eq(gpu_op(rb, "z", "sum", 3, tx), orc_op("z", "sum", 3, x))
    assert bits_eq(gpu_op(rb, "z", "nrm2", 3, tx), orc_op("z", "nrm2", 3, x))
