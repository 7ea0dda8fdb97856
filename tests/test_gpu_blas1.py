"""GPU parity of the BLAS1 path (sum, asum, nrm2, dot) against the oracle.

Bar: bit-identical doubles (NaN compared by class). Calls go through the C
ABI of libreproblas_hip.so, with host (staged) and device (in place)
operands. Reference behaviour: src/reproBLAS/rd{sum,asum,nrm2,dot}.c and the
binnedBLAS drivers; special-value classes per dmdsum.ccog:68-78.
"""
import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle

pytestmark = pytest.mark.gpu


def same(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return np.array([a]).view(np.uint64)[0] == np.array([b]).view(np.uint64)[0]


def orc_call(op, fold, x, y=None, inc=1):
    O = oracle()
    if op == "sum":
        return O.rdsum(fold, x, inc)
    if op == "asum":
        return O.rdasum(fold, x, inc)
    if op == "nrm2":
        return O.rdnrm2(fold, x, inc)
    return O.rddot(fold, x, y, inc, inc)


def gpu_call(rb, op, fold, x, y=None, inc=1):
    if op == "sum":
        return rb.rdsum(x, fold, inc)
    if op == "asum":
        return rb.rdasum(x, fold, inc)
    if op == "nrm2":
        return rb.rdnrm2(x, fold, inc)
    return rb.rddot(x, y, fold, inc, inc)


OPS = ["sum", "asum", "nrm2", "dot"]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("fold", [2, 3, 4, 6, 8])
@pytest.mark.parametrize("n", [1, 7, 511, 512, 513, 4095, 65537, 1 << 20])
def test_random_host_operands(gpu, op, fold, n):
    import reproblas_amd as rb

    x = dg.wide(n, 11 + n)
    y = dg.wide(n, 12 + n)
    r = gpu_call(rb, op, fold, x, y)
    e = orc_call(op, fold, x, y)
    assert same(r, e), (op, fold, n, r.hex(), e.hex())


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("dist", ["uniform", "wide", "ill", "big", "tiny", "mixed_range", "zeros", "neg_zeros",
                                  "constant"])
def test_distributions_device_operands(gpu, op, dist):
    import torch

    import reproblas_amd as rb

    n = 100003
    if dist == "uniform":
        x, y = dg.uniform(n, 1), dg.uniform(n, 2)
    elif dist == "wide":
        x, y = dg.wide(n, 1), dg.wide(n, 2)
    elif dist == "ill":
        x, y = dg.ill_conditioned(n, 1), dg.uniform(n, 2)
    else:
        x, y = dg.special(dist, n, 3), dg.special(dist, n, 4)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    for fold in (2, 3, 4):
        r = gpu_call(rb, op, fold, tx, ty)
        e = orc_call(op, fold, x, y)
        assert same(r, e), (op, dist, fold, r.hex(), e.hex())


# special fills x lengths; both_inf and inf_nan need two special members
SPECIAL_CASES = [(k, n) for k in ("nan", "pinf", "ninf", "both_inf", "inf_nan") for n in (1, 300, 70001)
                 if n >= 2 or k not in ("both_inf", "inf_nan")]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("kind,n", SPECIAL_CASES)
def test_special_values(gpu, op, kind, n):
    import reproblas_amd as rb

    x = dg.special(kind, n, 5)
    y = dg.uniform(n, 6)
    r = gpu_call(rb, op, 3, x, y)
    e = orc_call(op, 3, x, y)
    assert same(r, e), (op, kind, n, r, e)


@pytest.mark.parametrize("op", ["sum", "asum", "nrm2"])
@pytest.mark.parametrize("inc", [2, 3, -1, -2])
def test_strides(gpu, op, inc):
    import reproblas_amd as rb

    n = 20001
    base = dg.wide(n * abs(inc), 7)
    O = oracle()
    fn = {"sum": O.fn["rdsum"], "asum": O.fn["rdasum"], "nrm2": O.fn["rdnrm2"]}[op]
    import ctypes

    # negative increments: the pointer is the element with the lowest address
    # reached last (not adjusted), exactly as the reference indexes X[i*inc]
    ptr = base.ctypes.data + (0 if inc > 0 else (n - 1) * abs(inc) * 8)
    e = fn(3, n, ctypes.c_void_p(ptr), inc)
    L = rb.lib()
    g = {"sum": L.reproBLAS_rdsum, "asum": L.reproBLAS_rdasum, "nrm2": L.reproBLAS_rdnrm2}[op]
    rb._bind_stream()
    r = g(3, n, ctypes.c_void_p(ptr), inc)
    rb._check()
    assert same(r, e), (op, inc, r, e)


def test_dot_strided(gpu):
    import ctypes

    import reproblas_amd as rb

    n = 30001
    x = dg.wide(2 * n, 8)
    y = dg.wide(3 * n, 9)
    e = oracle().fn["rddot"](3, n, ctypes.c_void_p(x.ctypes.data), 2, ctypes.c_void_p(y.ctypes.data), 3)
    r = rb.rddot(x, y, 3, 2, 3, n=n)
    assert same(r, e)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ill_conditioned_shuffles_identical(gpu, seed):
    import reproblas_amd as rb

    n = 1 << 20
    base = rb.rdsum(dg.ill_conditioned(n, 0))
    r = rb.rdsum(dg.ill_conditioned(n, seed))
    assert same(r, base)
    assert same(r, oracle().rdsum(3, dg.ill_conditioned(n, seed)))


@pytest.mark.parametrize("op", ["sum", "asum", "dot"])
def test_accumulate_into_binned(gpu, op):
    """dbd* accumulate into an existing binned value (host and device Y)."""
    import torch

    import reproblas_amd as rb

    O = oracle()
    for fold in (2, 3, 4):
        x1, x2 = dg.wide(5000, 21), dg.wide(7000, 22) * 1e6
        y1, y2 = dg.wide(5000, 23), dg.wide(7000, 24)
        Yh = np.zeros(2 * fold)
        Ye = np.zeros(2 * fold)
        for xa, ya in ((x1, y1), (x2, y2)):
            if op == "sum":
                rb.dbdsum(Yh, xa, fold)
                O.dbdsum(fold, xa, Ye)
            elif op == "asum":
                rb.dbdasum(Yh, xa, fold)
                O.dbdasum(fold, xa, Ye)
            else:
                rb.dbddot(Yh, xa, ya, fold)
                O.dbddot(fold, xa, ya, Ye)
        assert np.array_equal(Yh.view(np.uint64), Ye.view(np.uint64)), (op, fold, Yh, Ye)
        # device binned value
        Yd = torch.zeros(2 * fold, dtype=torch.float64, device=gpu)
        for xa, ya in ((x1, y1), (x2, y2)):
            tx, ty = torch.from_numpy(xa).to(gpu), torch.from_numpy(ya).to(gpu)
            if op == "sum":
                rb.dbdsum(Yd, tx, fold)
            elif op == "asum":
                rb.dbdasum(Yd, tx, fold)
            else:
                rb.dbddot(Yd, tx, ty, fold)
        assert np.array_equal(Yd.cpu().numpy().view(np.uint64), Ye.view(np.uint64))


def test_ssq_accumulate_scale(gpu):
    import reproblas_amd as rb

    O = oracle()
    for fold in (2, 3):
        Yh, Ye = np.zeros(2 * fold), np.zeros(2 * fold)
        sh = se = 0.0
        for k, xa in enumerate((dg.wide(3000, 31), dg.wide(4000, 32) * 1e30, dg.wide(100, 33) * 1e-30)):
            sh = rb.dbdssq(Yh, xa, sh, fold)
            se = O.dbdssq(fold, xa, se, Ye)
            assert same(sh, se), (fold, k, sh, se)
            assert np.array_equal(Yh.view(np.uint64), Ye.view(np.uint64)), (fold, k, Yh, Ye)


def test_empty_and_negative_n(gpu):
    import reproblas_amd as rb

    x = np.zeros(0)
    assert rb.rdsum(x) == 0.0
    assert rb.rdnrm2(x) == 0.0
    assert rb.lib().reproBLAS_rdsum(3, -5, None, 1) == 0.0


def test_bad_fold_reports_error(gpu):
    import reproblas_amd as rb

    with pytest.raises(rb.ReproBLASError):
        rb.rdsum(np.ones(10), fold=1)


def test_device_result_api(gpu):
    """reproblas_hip_*_device: async, result in device memory."""
    import ctypes

    import torch

    import reproblas_amd as rb

    n = 1 << 18
    x, y = dg.wide(n, 41), dg.wide(n, 42)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    out = torch.zeros(4, dtype=torch.float64, device=gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = rb.lib()
    p = lambda t, off=0: ctypes.c_void_p(t.data_ptr() + 8 * off)
    assert L.reproblas_hip_rdsum_device(3, n, p(tx), 1, p(out, 0), s) == 0
    assert L.reproblas_hip_rdasum_device(3, n, p(tx), 1, p(out, 1), s) == 0
    assert L.reproblas_hip_rdnrm2_device(3, n, p(tx), 1, p(out, 2), s) == 0
    assert L.reproblas_hip_rddot_device(3, n, p(tx), 1, p(ty), 1, p(out, 3), s) == 0
    got = out.cpu().numpy()
    O = oracle()
    exp = [O.rdsum(3, x), O.rdasum(3, x), O.rdnrm2(3, x), O.rddot(3, x, y)]
    for g_, e_ in zip(got, exp):
        assert same(g_, e_)


def test_device_api_refuses_host_memory(gpu):
    """The *_device calls dereference their pointers in kernels; pageable host
    memory is refused on the host with an error (nothing is launched)."""
    import ctypes

    import torch

    import reproblas_amd as rb

    L = rb.lib()
    x = dg.wide(1000, 43)
    tx = torch.from_numpy(x).to(gpu)
    out = torch.zeros(2, dtype=torch.float64, device=gpu)
    host_out = np.zeros(2)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dp = lambda t: ctypes.c_void_p(t.data_ptr())
    hp = lambda a: ctypes.c_void_p(a.ctypes.data)
    assert L.reproblas_hip_rdsum_device(3, 1000, hp(x), 1, dp(out), s) != 0
    assert b"device-accessible" in L.reproblas_hip_last_error_string()
    assert L.reproblas_hip_rddot_device(3, 1000, dp(tx), 1, hp(x), 1, dp(out), s) != 0
    assert L.reproblas_hip_rdsum_device(3, 1000, dp(tx), 1, hp(host_out), s) != 0
    A = dg.normal(64 * 64, 44)
    tA = torch.from_numpy(A).to(gpu)
    C = torch.zeros(64 * 64, dtype=torch.float64, device=gpu)
    assert L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", 64, 64, 64, 1.0, hp(A), 64, dp(tA), 64, 0.0, dp(C),
                                         64, s) != 0
    assert L.reproblas_hip_rdgemv_device(3, b"C", b"N", 64, 64, 1.0, dp(tA), 64, hp(x), 1, 0.0, dp(C), 1, s) != 0
    rb.lib().reproblas_hip_clear_error()
    # the same calls on device memory still work afterwards
    assert L.reproblas_hip_rdsum_device(3, 1000, dp(tx), 1, dp(out), s) == 0
    assert same(out[0].item(), oracle().rdsum(3, x))


def test_checkpoint_resume(gpu, tmp_path):
    """A binned accumulator is a checkpoint: accumulate part of a stream on the
    device, save the state to disk, reload it (in a fresh buffer) and finish;
    the bits equal one uninterrupted call, on the device and host APIs."""
    import ctypes

    import torch

    import reproblas_amd as rb

    L = rb.lib()
    n = 300001
    x, y = dg.wide(n, 45), dg.wide(n, 46)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dp = lambda t, off=0: ctypes.c_void_p(t.data_ptr() + 8 * off)
    for fold in (2, 3, 5):
        Y = torch.zeros(2 * fold, dtype=torch.float64, device=gpu)
        assert L.reproblas_hip_dbddot_device(fold, 123457, dp(tx), 1, dp(ty), 1, dp(Y), s) == 0
        path = tmp_path / f"acc{fold}.npy"
        np.save(path, Y.cpu().numpy())
        del Y
        Y2 = torch.from_numpy(np.load(path)).to(gpu)
        assert L.reproblas_hip_dbddot_device(fold, n - 123457, dp(tx, 123457), 1, dp(ty, 123457), 1, dp(Y2), s) == 0
        one = torch.zeros(2 * fold, dtype=torch.float64, device=gpu)
        assert L.reproblas_hip_dbddot_device(fold, n, dp(tx), 1, dp(ty), 1, dp(one), s) == 0
        assert np.array_equal(Y2.cpu().numpy().view(np.uint64), one.cpu().numpy().view(np.uint64)), fold
        Ye = np.zeros(2 * fold)
        oracle().dbddot(fold, x, y, Ye)
        assert np.array_equal(one.cpu().numpy().view(np.uint64), Ye.view(np.uint64)), fold
        # host API: the same checkpoint taken through binnedBLAS_dbdsum
        H = np.zeros(2 * fold)
        rb.dbdsum(H, x[:777], fold)
        np.save(path, H)
        H2 = np.load(path)
        rb.dbdsum(H2, x[777:], fold)
        He = np.zeros(2 * fold)
        oracle().dbdsum(fold, x, He)
        assert np.array_equal(H2.view(np.uint64), He.view(np.uint64)), fold


def test_damax(gpu):
    import reproblas_amd as rb

    x = dg.wide(12345, 51)
    y = dg.wide(12345, 52)
    assert rb.damax(x) == np.max(np.abs(x))
    assert rb.damaxm(x, y) == np.max(np.abs(x * y))


@pytest.mark.slow
@pytest.mark.parametrize("fold", [2, 3, 4])
def test_full_size_dot_nrm2(gpu, fold):
    """Benchmark scale: N = 2^28 wide-range dot and nrm2, bit-identical to the oracle."""
    import torch

    import reproblas_amd as rb

    n = 1 << 28
    x = dg.wide(n, 101)
    y = dg.wide(n, 102)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    O = oracle()
    assert same(rb.rddot(tx, ty, fold), O.rddot(fold, x, y))
    assert same(rb.rdnrm2(tx, fold), O.rdnrm2(fold, x))
    if fold == 3:
        assert same(rb.rdsum(tx, fold), O.rdsum(fold, x))


@pytest.mark.slow
def test_largest_int_length(gpu):
    """N = 2^31 - 1, the largest length the reference's int arguments take
    (16 GiB per double operand): sum, asum, dot and nrm2 of constant vectors,
    whose binned sums are exact, so the results have closed forms (nrm2: the
    reference's scale * sqrt(conv), rdnrm2.c:31); float dot rounds n once
    (ssmconv sums in double)."""
    import math

    import torch

    import reproblas_amd as rb

    n = (1 << 31) - 1
    O = oracle()
    x = torch.full((n,), 0.5, device=gpu, dtype=torch.float64)
    assert rb.rdsum(x) == 0.5 * n
    y = torch.full((n,), 3.0, device=gpu, dtype=torch.float64)
    assert rb.rddot(x, y) == 1.5 * n
    del y
    x.fill_(-1.0)
    assert rb.rdasum(x) == float(n)
    sc = O.dscale(1.0)
    assert same(rb.rdnrm2(x), sc * math.sqrt(float(n) / (sc * sc)))
    del x
    torch.cuda.empty_cache()
    xs = torch.ones(n, device=gpu, dtype=torch.float32)
    assert rb.rsdot(xs, xs) == np.float32(float(n))
    del xs
    torch.cuda.empty_cache()


@pytest.mark.slow
def test_full_size_ill_conditioned_sum(gpu):
    """N = 2^28 ill-conditioned sum: N/4 values x, their N/4 negations and N/2
    terms of 1e-12*U(-1,1) noise, in three random orders. The three sums are
    bit-identical to each other and to the oracle (order independence at full
    size)."""
    import torch

    import reproblas_amd as rb

    n = 1 << 28
    g = torch.Generator(device=gpu)
    g.manual_seed(77)
    q = n // 4
    v = (torch.rand(q, generator=g, device=gpu, dtype=torch.float64) + 1.0) * torch.ldexp(
        torch.ones(q, device=gpu, dtype=torch.float64),
        torch.randint(-40, 41, (q,), generator=g, device=gpu, dtype=torch.int32))
    noise = (torch.rand(n - 2 * q, generator=g, device=gpu, dtype=torch.float64) * 2 - 1) * 1e-12
    base = torch.cat([v, -v, noise])
    del v, noise
    sums = []
    for seed in (1, 2, 3):
        g.manual_seed(seed)
        perm = torch.randperm(n, generator=g, device=gpu)
        xs = base[perm]
        del perm
        sums.append(rb.rdsum(xs))
        if seed == 1:
            host = xs.cpu().numpy()
        del xs
    assert same(sums[0], sums[1]) and same(sums[0], sums[2])
    assert same(sums[0], oracle().rdsum(3, host))


@pytest.mark.parametrize("fold", [9, 13, 26, 52])
def test_high_folds(gpu, fold):
    """Folds above 8 run the general-capacity kernels (slots in private memory)."""
    import torch

    import reproblas_amd as rb

    O = oracle()
    for n in (1, 777, 70001):
        x, y = dg.wide(n, 300 + n), dg.wide(n, 400 + n)
        for op in OPS:
            assert same(gpu_call(rb, op, fold, x, y), orc_call(op, fold, x, y)), (fold, n, op)
    x = dg.special("mixed_range", 5000, 9)
    assert same(rb.rdsum(x, fold), O.rdsum(fold, x))
    Y, Ye = np.zeros(2 * fold), np.zeros(2 * fold)
    rb.dbdsum(Y, x, fold)
    O.dbdsum(fold, x, Ye)
    assert np.array_equal(Y.view(np.uint64), Ye.view(np.uint64))
    tx = torch.from_numpy(x).to(gpu)
    assert same(rb.rdnrm2(tx, fold), O.rdnrm2(fold, x))
