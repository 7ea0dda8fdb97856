"""Benchmark of the binned FP64 path on one MI355X.

Headline (the JSON line's `value`): reproducible (bit-exact) ddot, fold 3,
N = 2^28 FP64 with wide dynamic range x = s*m*2^e (s = +-1, m ~ U[1,2),
e ~ U{-40..40}), both operands resident in HBM; GB/s of algorithmic traffic
(16 bytes per term). A "step" is one ddot over the whole vector pair.

Also reported under "extra": dsum 2^20 (uniform), dsum / dnrm2 2^28, ddot and
dnrm2 2^28 at folds 2, 3, 4 and dsum with incX = 2, dgemv 16384^2 (N and T),
dgemm 4096^3 (GFLOP/s and fraction of the FP64 VALU issue rate), the ragged
batch and the MLP step, the single-precision path (sdot / ssum / snrm2
2^28 floats, sgemv 16384^2, sgemm 4096^3) and the complex path (zdotc / zsum /
dznrm2 over 2^27 double complex, cdotc over 2^27 float complex, zgemv
8192^2 N and T, zgemm 2048^3), each as one timed configuration.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# FP64 VALU issue rate: 256 CUs x 64 lanes x 2.4 GHz (one op per lane per
# clock; the 78.6 TFLOP/s figure counts an FMA as two)
VALU_PEAK_OPS = 256 * 64 * 2.4e9
# FP32 VALU issue rate: 256 CUs x 128 lanes x 2.4 GHz (the published 157.3
# TFLOP/s FP32 vector figure, an FMA counted as two)
VALU_PEAK_OPS_FP32 = 256 * 128 * 2.4e9
METRIC = "bit-exact ddot GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def wide_gpu(n, seed, device):
    """x = s*m*2^e on the GPU from a seeded generator."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    s = torch.randint(0, 2, (n,), generator=g, device=device, dtype=torch.int8).to(torch.float64).mul_(2).sub_(1)
    m = torch.rand(n, generator=g, device=device, dtype=torch.float64).add_(1.0)
    e = torch.randint(-40, 41, (n,), generator=g, device=device, dtype=torch.int32)
    return torch.ldexp(m, e).mul_(s)


# untimed calls before each HBM-bound extra configuration (a few ms: the
# clocks settle before the timed calls)
WARM_HBM = 20


def time_launches(fn, steps, warmup, stream):
    """Average time per call of fn (which enqueues on `stream`), HIP events."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / 1e3 / steps


def bits(v):
    return int(np.array([v]).view(np.uint64)[0])


def cpu_baseline(x_sample, y_sample):
    """The reference's own ddot (oracle/_ref, compiled from its sources) on the
    host, single thread, on a bounded sample; falls back to the oracle port."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cpu_libs import oracle, reference

    lib = reference("avx")
    kind = "reference"
    if lib is None:
        lib, kind = oracle(), "port"
    n = len(x_sample)
    reps, t_total, res = 0, 0.0, None
    while t_total < 10.0 and reps < 5000:
        t = time.perf_counter()
        res = lib.rddot(3, x_sample, y_sample)
        t_total += time.perf_counter() - t
        reps += 1
    gbs = 16.0 * n * reps / t_total / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"rddot fold 3 on the first {n} terms of the same vectors, {reps} repetitions, "
                      f"{t_total:.1f} s"}, res


def cpu_baseline_parallel(x_sample, y_sample, ref_value):
    """The same sample sharded over the host's cores: each thread runs the
    reference's binnedBLAS_dbddot on its shard (ctypes releases the GIL), the
    binned shards merge in order with binned_dbdbadd and binned_ddbconv
    converts; the bits equal the single-thread rddot."""
    import concurrent.futures

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cpu_libs import oracle, reference

    lib = reference("avx")
    kind = "reference"
    if lib is None:
        lib, kind = oracle(), "port"
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    n = len(x_sample)
    bounds = [n * i // threads for i in range(threads + 1)]
    parts = [np.zeros(6) for _ in range(threads)]
    dbddot = lib.fn["dbddot"]
    vpa = lambda a: ctypes.c_void_p(a.ctypes.data)

    def shard(i):
        parts[i][:] = 0.0
        lo, hi = bounds[i], bounds[i + 1]
        dbddot(3, hi - lo, ctypes.c_void_p(x_sample.ctypes.data + 8 * lo), 1,
               ctypes.c_void_p(y_sample.ctypes.data + 8 * lo), 1, vpa(parts[i]))

    reps, t_total, res = 0, 0.0, None
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while t_total < 5.0 and reps < 20000:
            t = time.perf_counter()
            list(ex.map(shard, range(threads)))
            acc = parts[0].copy()
            for i in range(1, threads):
                lib.fn["dbdbadd"](3, vpa(parts[i]), vpa(acc))
            res = lib.fn["ddbconv"](3, vpa(acc))
            t_total += time.perf_counter() - t
            reps += 1
    gbs = 16.0 * n * reps / t_total / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"the same {n}-term sample in {threads} shards (dbddot per thread, dbdbadd merge), "
                      f"{reps} repetitions, {t_total:.1f} s",
            "bits_equal_single_thread": bits(res) == bits(ref_value)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--quick", action="store_true", help="headline only (no extra configurations)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--only", choices=["dsum_2p20", "dsum_2p28", "dnrm2_2p28", "folds", "dgemv", "dgemm", "ragged_batch",
                                       "mlp", "single", "complex"],
                    help="run one extra configuration alone and print its record (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain device copy and a plain torch sum of 2^28 doubles (scales the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the multi-GPU reduction is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import torch

    import reproblas_amd as rb

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        calibrate(dev, args)
        return
    L = rb.lib()
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    if args.only:
        print(json.dumps({args.only: extras(L, rb, dev, stream, sp, vp, args, only=args.only)[args.only]}), flush=True)
        return

    # ------------------------------------------------------------ headline --
    n = 1 << 28
    log(f"[bench] generating ddot operands N={n}")
    x = wide_gpu(n, 1, dev)
    y = wide_gpu(n, 2, dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)

    def ddot():
        L.reproblas_hip_rddot_device(3, n, vp(x), 1, vp(y), 1, vp(out), sp)

    log("[bench] timing ddot")
    t = time_launches(ddot, args.steps, args.warmup, stream)
    rb._check()
    value = 16.0 * n / t / 1e9
    log(f"[bench] ddot {t * 1e3:.3f} ms/call, {value:.1f} GB/s")

    # bit-exactness spot check on a sample against the CPU reference
    cpu = cpu_par = None
    ref_sample = gpu_sample = None
    if not args.no_cpu:
        ns = 1 << 24
        xs = x[:ns].cpu().numpy()
        ys = y[:ns].cpu().numpy()
        L.reproblas_hip_rddot_device(3, ns, vp(x), 1, vp(y), 1, vp(out[1:]), sp)
        torch.cuda.synchronize()
        gpu_sample = float(out[1].item())
        log("[bench] CPU baseline")
        cpu, ref_sample = cpu_baseline(xs, ys)
        cpu_par = cpu_baseline_parallel(xs, ys, ref_sample)
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_blas1_ddot_2p28.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "reproBLAS_ddot fold=3 N=2^28 wide-range x=s*m*2^e, e in [-40,40], incX=incY=1",
                   "fold": 3, "n": n, "bytes_per_step": 16 * n},
        "roofline": {"bound": "hbm", "achieved": round(value, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(value / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_par,
        "parity_sample": None if ref_sample is None else (bits(ref_sample) == bits(gpu_sample)),
    }
    del x, y
    torch.cuda.empty_cache()

    if not args.quick:
        line["extra"] = extras(L, rb, dev, stream, sp, vp, args)
    print(json.dumps(line), flush=True)


def calibrate(dev, args):
    """Reference kernels of known traffic: copy (8+8 bytes per element) and a
    plain sum (8 bytes per element) of 2^28 doubles."""
    import torch

    n = 1 << 28
    x = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)
    t_copy = time_launches(lambda: y.copy_(x), args.steps, args.warmup, stream)
    t_sum = time_launches(lambda: x.sum(), args.steps, args.warmup, stream)
    print(json.dumps({"copy_bytes_per_launch": 16 * n, "copy_ms": round(t_copy * 1e3, 4),
                      "copy_GB/s": round(16 * n / t_copy / 1e9, 1), "sum_bytes_per_launch": 8 * n,
                      "sum_ms": round(t_sum * 1e3, 4), "sum_GB/s": round(8 * n / t_sum / 1e9, 1)}), flush=True)


def extras(L, rb, dev, stream, sp, vp, args, only=None):
    import torch

    ex = {}
    want = lambda name: only is None or only == name
    steps = max(3, args.steps // 2)
    out = torch.zeros(8, dtype=torch.float64, device=dev)

    g = torch.Generator(device=dev)
    g.manual_seed(0)
    if want("dsum_2p20"):
        # dsum fold 3, N = 2^20 uniform(-1,1): launch/latency bound, report us
        x = torch.rand(1 << 20, generator=g, device=dev, dtype=torch.float64).mul_(2).sub_(1)
        t = time_launches(lambda: L.reproblas_hip_rdsum_device(3, 1 << 20, vp(x), 1, vp(out), sp), 50, 5, stream)
        ex["dsum_2p20"] = {"us_per_call": round(t * 1e6, 2), "GB/s": round(8 * (1 << 20) / t / 1e9, 1)}
        log(f"[bench] dsum 2^20: {t * 1e6:.1f} us")

    if want("dsum_2p28") or want("dnrm2_2p28"):
        n = 1 << 28
        x = wide_gpu(n, 3, dev)
        for name, fn in (("dsum_2p28", lambda: L.reproblas_hip_rdsum_device(3, n, vp(x), 1, vp(out), sp)),
                         ("dnrm2_2p28", lambda: L.reproblas_hip_rdnrm2_device(3, n, vp(x), 1, vp(out), sp))):
            if not want(name):
                continue
            t = time_launches(fn, steps, WARM_HBM, stream)
            gbs = 8 * n / t / 1e9
            ex[name] = {"ms": round(t * 1e3, 3), "GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
            log(f"[bench] {name}: {t * 1e3:.3f} ms, {gbs:.1f} GB/s")
        del x
        torch.cuda.empty_cache()

    if want("folds"):
        # ddot and dnrm2 at folds 2, 3, 4 (N = 2^28 wide-range), and dsum with incX = 2
        n = 1 << 28
        x = wide_gpu(n, 1, dev)
        y = wide_gpu(n, 2, dev)
        rec = {"ddot": {}, "dnrm2": {}}
        for fold in (2, 3, 4):
            t = time_launches(lambda: L.reproblas_hip_rddot_device(fold, n, vp(x), 1, vp(y), 1, vp(out), sp),
                              steps, WARM_HBM, stream)
            rec["ddot"][str(fold)] = {"ms": round(t * 1e3, 3), "GB/s": round(16 * n / t / 1e9, 1)}
            t = time_launches(lambda: L.reproblas_hip_rdnrm2_device(fold, n, vp(x), 1, vp(out), sp), steps, WARM_HBM, stream)
            rec["dnrm2"][str(fold)] = {"ms": round(t * 1e3, 3), "GB/s": round(8 * n / t / 1e9, 1)}
        h = n // 2
        t = time_launches(lambda: L.reproblas_hip_rdsum_device(3, h, vp(x), 2, vp(out), sp), steps, WARM_HBM, stream)
        rec["dsum_inc2"] = {"terms": h, "ms": round(t * 1e3, 3), "GB/s_of_terms": round(8 * h / t / 1e9, 1),
                            "GB/s_of_span": round(16 * h / t / 1e9, 1)}
        ex["folds"] = ex["folds_2p28"] = rec
        log(f"[bench] folds: {rec}")
        del x, y
        torch.cuda.empty_cache()

    if want("dgemv"):
        # dgemv 16384^2 column-major, fold 3, alpha 1, beta 0.5
        m = 16384
        A = torch.randn(m * m, generator=g, device=dev, dtype=torch.float64)
        xv = torch.randn(m, generator=g, device=dev, dtype=torch.float64)
        yv = torch.randn(m, generator=g, device=dev, dtype=torch.float64)
        for tr in ("N", "T"):
            fn = lambda: L.reproblas_hip_rdgemv_device(3, b"C", tr.encode(), m, m, 1.0, vp(A), m, vp(xv), 1, 0.5,
                                                       vp(yv), 1, sp)
            t = time_launches(fn, steps, WARM_HBM, stream)
            byts = 8 * (m * m + m + 2 * m)
            ex[f"dgemv_{tr}_16384"] = {"ms": round(t * 1e3, 3), "GB/s": round(byts / t / 1e9, 1),
                                       "hbm_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"[bench] dgemv {tr}: {t * 1e3:.3f} ms, {byts / t / 1e9:.1f} GB/s")
        ex["dgemv"] = {"N": ex["dgemv_N_16384"], "T": ex["dgemv_T_16384"]}
        del A
        torch.cuda.empty_cache()

    if want("dgemm"):
        # dgemm 4096^3 column-major NN, fold 3, alpha 1, beta 0
        k = 4096
        A = torch.randn(k * k, generator=g, device=dev, dtype=torch.float64)
        B = torch.randn(k * k, generator=g, device=dev, dtype=torch.float64)
        C = torch.zeros(k * k, device=dev, dtype=torch.float64)
        fn = lambda: L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", k, k, k, 1.0, vp(A), k, vp(B), k, 0.0, vp(C),
                                                   k, sp)
        t = time_launches(fn, 3, 2, stream)
        ops = 11.0 * k ** 3  # (4*fold-1) ops per product, the reference's model (bench_rdgemm.c:125-127)
        ex["dgemm"] = ex["dgemm_NN_4096"] = {"ms": round(t * 1e3, 2), "GFLOP/s": round(ops / t / 1e9, 1),
                                             "valu_frac": round(ops / t / VALU_PEAK_OPS, 4),
                                             "products_per_s": round(k ** 3 / t, 1)}
        log(f"[bench] dgemm: {t * 1e3:.2f} ms, {ops / t / 1e9:.1f} GFLOP/s (model ops), "
            f"{ops / t / VALU_PEAK_OPS:.3f} of VALU")
        del A, B, C
        torch.cuda.empty_cache()
    if want("single"):
        ex["single"] = single(L, dev, stream, sp, vp, steps)
    if want("complex"):
        ex["complex"] = complex_path(L, dev, stream, sp, vp, steps)
    if want("ragged_batch"):
        ex["ragged_batch"] = ragged_batch(L, dev, stream, sp, vp)
    if want("mlp"):
        ex["mlp"] = ex["mlp_fwd_bwd"] = mlp(dev, stream, args)
    if only is None:
        ex.pop("folds", None)
        ex.pop("dgemv", None)
        ex.pop("dgemm", None)
        ex.pop("mlp", None)
    return ex


def single(L, dev, stream, sp, vp, steps):
    """The float path at fold 3: sdot / ssum / snrm2 over N = 2^28 wide-range
    floats (x = s*m*2^e rounded to float, e in [-40,40]), sgemv 16384^2 (N and
    T, beta 0.5) and sgemm 4096^3 NN."""
    import torch

    rec = {}
    n = 1 << 28
    x = wide_gpu(n, 11, dev).float()
    y = wide_gpu(n, 12, dev).float()
    torch.cuda.empty_cache()
    out = torch.zeros(8, dtype=torch.float32, device=dev)
    for name, fn, byts in (
            ("sdot", lambda: L.reproblas_hip_rsdot_device(3, n, vp(x), 1, vp(y), 1, vp(out), sp), 8 * n),
            ("ssum", lambda: L.reproblas_hip_rssum_device(3, n, vp(x), 1, vp(out), sp), 4 * n),
            ("snrm2", lambda: L.reproblas_hip_rsnrm2_device(3, n, vp(x), 1, vp(out), sp), 4 * n)):
        t = time_launches(fn, max(steps, 10), 20, stream)
        rec[name + "_2p28"] = {"ms": round(t * 1e3, 3), "GB/s": round(byts / t / 1e9, 1),
                               "hbm_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] {name} 2^28: {t * 1e3:.3f} ms, {byts / t / 1e9:.1f} GB/s")
    del x, y
    torch.cuda.empty_cache()
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    m = 16384
    A = torch.randn(m * m, generator=g, device=dev, dtype=torch.float32)
    xv = torch.randn(m, generator=g, device=dev, dtype=torch.float32)
    yv = torch.randn(m, generator=g, device=dev, dtype=torch.float32)
    for tr in ("N", "T"):
        fn = lambda: L.reproblas_hip_rsgemv_device(3, b"C", tr.encode(), m, m, 1.0, vp(A), m, vp(xv), 1, 0.5,
                                                   vp(yv), 1, sp)
        t = time_launches(fn, max(steps, 10), 10, stream)
        byts = 4 * (m * m + 3 * m)
        rec[f"sgemv_{tr}_16384"] = {"ms": round(t * 1e3, 3), "GB/s": round(byts / t / 1e9, 1),
                                    "hbm_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] sgemv {tr}: {t * 1e3:.3f} ms, {byts / t / 1e9:.1f} GB/s")
    del A
    torch.cuda.empty_cache()
    k = 4096
    A = torch.randn(k * k, generator=g, device=dev, dtype=torch.float32)
    B = torch.randn(k * k, generator=g, device=dev, dtype=torch.float32)
    C = torch.zeros(k * k, device=dev, dtype=torch.float32)
    fn = lambda: L.reproblas_hip_rsgemm_device(3, b"C", b"N", b"N", k, k, k, 1.0, vp(A), k, vp(B), k, 0.0, vp(C), k,
                                               sp)
    t = time_launches(fn, 3, 2, stream)
    ops = 11.0 * k ** 3
    rec["sgemm_NN_4096"] = {"ms": round(t * 1e3, 2), "GFLOP/s": round(ops / t / 1e9, 1),
                            "valu_frac": round(ops / t / VALU_PEAK_OPS_FP32, 4)}
    log(f"[bench] sgemm: {t * 1e3:.2f} ms, {ops / t / 1e9:.1f} GFLOP/s (model ops)")
    del A, B, C
    torch.cuda.empty_cache()
    import reproblas_amd as rb

    rb._check()
    return rec


def complex_path(L, dev, stream, sp, vp, steps):
    """Complex binned reductions at fold 3: zdotc, zsum, dznrm2 over N = 2^27
    double complex (wide-range parts), cdotc over 2^27 float complex, zgemv
    8192^2 column-major N (beta 0.5) and zgemm 2048^3 column-major NN."""
    import torch

    rec = {}
    n = 1 << 27
    x = wide_gpu(2 * n, 21, dev)  # interleaved (re, im)
    y = wide_gpu(2 * n, 22, dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    for name, fn, byts in (
            ("zdotc", lambda: L.reproblas_hip_rzdotc_device(3, n, vp(x), 1, vp(y), 1, vp(out), sp), 32 * n),
            ("zsum", lambda: L.reproblas_hip_rzsum_device(3, n, vp(x), 1, vp(out), sp), 16 * n),
            ("dznrm2", lambda: L.reproblas_hip_rdznrm2_device(3, n, vp(x), 1, vp(out), sp), 16 * n)):
        t = time_launches(fn, max(steps, 10), 10, stream)
        rec[name + "_2p27"] = {"ms": round(t * 1e3, 3), "GB/s": round(byts / t / 1e9, 1),
                               "hbm_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] {name} 2^27: {t * 1e3:.3f} ms, {byts / t / 1e9:.1f} GB/s")
    xf, yf = x.float(), y.float()
    del x, y
    torch.cuda.empty_cache()
    outf = torch.zeros(8, dtype=torch.float32, device=dev)
    t = time_launches(lambda: L.reproblas_hip_rcdotc_device(3, n, vp(xf), 1, vp(yf), 1, vp(outf), sp), max(steps, 10),
                      20, stream)
    rec["cdotc_2p27"] = {"ms": round(t * 1e3, 3), "GB/s": round(16 * n / t / 1e9, 1),
                         "hbm_frac": round(16 * n / t / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"[bench] cdotc 2^27: {t * 1e3:.3f} ms")
    del xf, yf
    torch.cuda.empty_cache()
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    al = np.array([1.0, 0.0])
    be = np.array([0.5, 0.0])
    pal, pbe = ctypes.c_void_p(al.ctypes.data), ctypes.c_void_p(be.ctypes.data)
    m = 8192
    A = torch.randn(2 * m * m, generator=g, device=dev, dtype=torch.float64)
    xv = torch.randn(2 * m, generator=g, device=dev, dtype=torch.float64)
    yv = torch.randn(2 * m, generator=g, device=dev, dtype=torch.float64)
    for tr in ("N", "T"):
        fn = lambda: L.reproblas_hip_rzgemv_device(3, b"C", tr.encode(), m, m, pal, vp(A), m, vp(xv), 1, pbe, vp(yv),
                                                   1, sp)
        t = time_launches(fn, max(steps, 10), 5, stream)
        byts = 16 * (m * m + 3 * m)
        rec[f"zgemv_{tr}_8192"] = {"ms": round(t * 1e3, 3), "GB/s": round(byts / t / 1e9, 1),
                                   "hbm_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] zgemv {tr} 8192: {t * 1e3:.3f} ms")
    del A
    torch.cuda.empty_cache()
    k = 2048
    A = torch.randn(2 * k * k, generator=g, device=dev, dtype=torch.float64)
    B = torch.randn(2 * k * k, generator=g, device=dev, dtype=torch.float64)
    C = torch.zeros(2 * k * k, device=dev, dtype=torch.float64)
    bz = np.array([0.0, 0.0])
    pbz = ctypes.c_void_p(bz.ctypes.data)
    fn = lambda: L.reproblas_hip_rzgemm_device(3, b"C", b"N", b"N", k, k, k, pal, vp(A), k, vp(B), k, pbz, vp(C), k,
                                               sp)
    t = time_launches(fn, 3, 2, stream)
    ops = 4 * 11.0 * k ** 3  # four real products per complex product
    rec["zgemm_NN_2048"] = {"ms": round(t * 1e3, 2), "GFLOP/s": round(ops / t / 1e9, 1),
                            "valu_frac": round(ops / t / VALU_PEAK_OPS, 4)}
    log(f"[bench] zgemm 2048: {t * 1e3:.2f} ms")
    del A, B, C
    torch.cuda.empty_cache()
    import reproblas_amd as rb

    rb._check()
    return rec


def ragged_batch(L, dev, stream, sp, vp):
    """10,000 dot segments, lengths log-uniform in [16, 2^20] (CSR), wide-range
    values, plus 512 column-major NN gemms with M, N, K uniform in
    {32, 64, ..., 512}; fold 3. One graph launch of the batch plan against the
    same work issued as per-call launches (each its own reproBLAS call)."""
    import torch

    from reproblas_amd.autotune import tune_gemm_batch
    from reproblas_amd.batch import Batch

    g = np.random.Generator(np.random.PCG64(7))
    nseg = 10000
    lens = np.exp(g.uniform(np.log(16), np.log(1 << 20), nseg)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(off[-1])
    x = wide_gpu(total, 11, dev)
    y = wide_gpu(total, 12, dev)
    res = torch.zeros(nseg, dtype=torch.float64, device=dev)
    res2 = torch.zeros(nseg, dtype=torch.float64, device=dev)
    sizes = g.integers(1, 17, (512, 3)) * 32
    gt = torch.Generator(device=dev)
    gt.manual_seed(5)
    triples = [(torch.randn(int(K), int(M), generator=gt, device=dev, dtype=torch.float64),
                torch.randn(int(N), int(K), generator=gt, device=dev, dtype=torch.float64),
                torch.zeros(int(N), int(M), device=dev, dtype=torch.float64)) for M, N, K in sizes]
    log(f"[bench] ragged batch: {nseg} dot segments, {total} terms; 512 gemms, {int(np.prod(sizes, axis=1).sum())} products")
    tuned = tune_gemm_batch("N", "N", triples, use_cache=False, reps=2)
    b = Batch(3)
    b.add_segments("dot", off, x, y, res)
    b.add_gemms("N", "N", triples, tiling=tuned["tiling"])
    t_graph = time_launches(lambda: b.launch(graph=True), 3, 1, stream)

    xp, yp, rp = x.data_ptr(), y.data_ptr(), res2.data_ptr()

    def per_call():
        for s in range(nseg):
            L.reproblas_hip_rddot_device(3, int(lens[s]), ctypes.c_void_p(xp + 8 * int(off[s])), 1,
                                         ctypes.c_void_p(yp + 8 * int(off[s])), 1, ctypes.c_void_p(rp + 8 * s), sp)
        for (M, N, K), (A, B, C) in zip(sizes, triples):
            L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", int(M), int(N), int(K), 1.0, vp(A), int(M), vp(B),
                                          int(K), 0.0, vp(C), int(M), sp)

    t_calls = time_launches(per_call, 1, 1, stream)
    same = bool(torch.equal(res.view(torch.int64), res2.view(torch.int64)))
    b.close()
    log(f"[bench] ragged batch: graph {t_graph * 1e3:.2f} ms, per-call {t_calls * 1e3:.2f} ms, "
        f"speedup {t_calls / t_graph:.2f}x, identical={same}")
    return {"graph_ms": round(t_graph * 1e3, 3), "per_call_ms": round(t_calls * 1e3, 3),
            "speedup": round(t_calls / t_graph, 2), "segments": nseg, "segment_terms": total, "gemms": 512,
            "gemm_products": int(np.prod(sizes, axis=1).sum()), "tiling": tuned["tiling"],
            "results_identical": same}


def mlp(dev, stream, args):
    """3-layer FP64 MLP 1024-4096-4096-1024 (tanh), batch 2048, MSE loss:
    one forward + backward, every matmul binned (fold 3)."""
    import torch

    import reproblas_amd.autograd as ag

    g = torch.Generator(device=dev)
    g.manual_seed(3)
    model = ag.ReproMLP((1024, 4096, 4096, 1024), generator=g, device=dev)
    xb = torch.randn(2048, 1024, generator=g, device=dev, dtype=torch.float64)
    tb = torch.randn(2048, 1024, generator=g, device=dev, dtype=torch.float64)
    t = time_launches(lambda: ag.mlp_step(model, xb, tb), 2, 1, stream)
    fwd = 2048 * (1024 * 4096 + 4096 * 4096 + 4096 * 1024)
    products = fwd + 2048 * (1024 * 4096 + 4096 * 4096 + 4096 * 1024) + 2048 * (4096 * 4096 + 4096 * 1024)
    log(f"[bench] mlp fwd+bwd: {t * 1e3:.1f} ms, {products / t / 1e12:.2f} T products/s")
    return {"ms": round(t * 1e3, 2), "products_per_step": products,
            "valu_frac": round(11.0 * products / t / VALU_PEAK_OPS, 4)}


if __name__ == "__main__":
    main()
