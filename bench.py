"""Benchmark of the binned FP64 path on one MI355X.

Headline (the JSON line's `value`): reproducible (bit-exact) ddot, fold 3,
N = 2^28 FP64 with wide dynamic range x = s*m*2^e (s = +-1, m ~ U[1,2),
e ~ U{-40..40}), both operands resident in HBM; GB/s of algorithmic traffic
(16 bytes per term). A "step" is one ddot over the whole vector pair.

Also reported under "extra", each from per-call HIP events (median, minimum
and mean of at least 20 calls after warm-up; 10 for the MLP step): dsum 2^20
(uniform), dsum / dnrm2 2^28, ddot and dnrm2 2^28 at folds 2, 3, 4 and dsum
with incX = 2, dgemv 16384^2 (N and T, beta 0 and 0.5), dgemm 4096^3 (NN with
beta 0 and 0.5, TN; GFLOP/s and fraction of the FP64 VALU issue rate), the
ragged batch (autotuned tiling and chunk; graph launch against per-call
launches, every output compared) and the MLP step, the single-precision path
(sdot / ssum / snrm2 2^28 floats, sgemv 16384^2, sgemm 4096^3) and the
complex path (zdotc / zsum / dznrm2 over 2^27 double complex, cdotc over 2^27
float complex, zgemv 8192^2 N and T, zgemm 2048^3).

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


def time_each(fn, n, warmup, stream):
    """Per-call times (seconds) of n calls of fn, each between its own pair of
    HIP events on `stream`, after `warmup` untimed calls."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    for e0, e1 in ev:
        e0.record(stream)
        fn()
        e1.record(stream)
    torch.cuda.synchronize()
    return [e0.elapsed_time(e1) / 1e3 for e0, e1 in ev]


def median(v):
    return float(np.median(np.asarray(v)))


def stats(ts):
    """Median, minimum and mean of per-call times, in ms."""
    a = np.asarray(ts) * 1e3
    return {"median_ms": round(float(np.median(a)), 4), "min_ms": round(float(a.min()), 4),
            "mean_ms": round(float(a.mean()), 4), "calls": len(a)}


def bits(v):
    return int(np.array([v]).view(np.uint64)[0])


def host_cpu():
    """CPU model and the number of CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return model, ncpu


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
    model, ncpu = host_cpu()
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": kind, "cpu_model": model,
            "host_cpus": ncpu,
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
    model, ncpu = host_cpu()
    # every CPU this process may use, up to the host's thread budget
    # (OMP_NUM_THREADS, which the GPU hosts set to their per-GPU CPU share)
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(ncpu, budget) if budget > 0 else ncpu)
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
            "host_cpus": ncpu, "cpu_model": model, "omp_num_threads": budget or None,
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
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the library runs the reduction on one GPU; the exchange between processes is reproblas_amd.dist)", file=sys.stderr)
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
    # the same call, each of 20 launches between its own events
    per_call = stats(time_each(ddot, 20, 0, stream))
    per_call["GB/s"] = round(16.0 * n / (per_call["median_ms"] / 1e3) / 1e9, 1)
    # the vendor's (non-reproducible) fp64 dot on the same operands: the rate
    # a plain streaming read reaches on this GPU, for comparison only
    t_vendor = time_launches(lambda: torch.dot(x, y), args.steps, args.warmup, stream)
    vendor = {"kernel": "torch.dot (rocBLAS ddot, not reproducible)", "ms_per_call": round(t_vendor * 1e3, 4),
              "GB/s": round(16.0 * n / t_vendor / 1e9, 1), "throughput_ratio_ours_to_vendor": round(t_vendor / t, 4)}
    log(f"[bench] vendor ddot {t_vendor * 1e3:.3f} ms/call, {vendor['GB/s']:.1f} GB/s")

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
        "per_call": per_call,
        "vendor_ddot": vendor,
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


def hbm_rec(ts, byts):
    """Per-call statistics of an HBM-bound configuration, GB/s from the median."""
    r = stats(ts)
    gbs = byts / (r["median_ms"] / 1e3) / 1e9
    r.update({"GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
    return r


def valu_rec(ts, ops, peak=VALU_PEAK_OPS):
    """Per-call statistics of a VALU-bound configuration (model ops), from the median."""
    r = stats(ts)
    t = r["median_ms"] / 1e3
    r.update({"GFLOP/s": round(ops / t / 1e9, 1), "valu_frac": round(ops / t / peak, 4),
              "frac_of_published_fp64": round(ops / t / 78.6e12, 4)})
    return r


def extras(L, rb, dev, stream, sp, vp, args, only=None):
    """Every extra configuration: per-call HIP-event times, median of at least
    20 calls after warm-up (10 for the MLP step)."""
    import torch

    ex = {}
    want = lambda name: only is None or only == name
    calls = max(20, args.steps)
    out = torch.zeros(8, dtype=torch.float64, device=dev)

    g = torch.Generator(device=dev)
    g.manual_seed(0)
    if want("dsum_2p20"):
        # dsum fold 3, N = 2^20 uniform(-1,1): launch/latency bound, report us
        x = torch.rand(1 << 20, generator=g, device=dev, dtype=torch.float64).mul_(2).sub_(1)
        fn = lambda: L.reproblas_hip_rdsum_device(3, 1 << 20, vp(x), 1, vp(out), sp)
        ts = time_each(fn, 100, 20, stream)
        t_avg = time_launches(fn, 100, 5, stream)
        r = hbm_rec(ts, 8 * (1 << 20))
        r["us_per_call"] = round(r["median_ms"] * 1e3, 2)
        r["back_to_back_us_per_call"] = round(t_avg * 1e6, 2)
        ex["dsum_2p20"] = r
        log(f"[bench] dsum 2^20: {r['us_per_call']:.2f} us median, {t_avg * 1e6:.2f} us back to back")

    if want("dsum_2p28") or want("dnrm2_2p28"):
        n = 1 << 28
        x = wide_gpu(n, 3, dev)
        for name, fn in (("dsum_2p28", lambda: L.reproblas_hip_rdsum_device(3, n, vp(x), 1, vp(out), sp)),
                         ("dnrm2_2p28", lambda: L.reproblas_hip_rdnrm2_device(3, n, vp(x), 1, vp(out), sp))):
            if not want(name):
                continue
            ex[name] = hbm_rec(time_each(fn, calls, WARM_HBM, stream), 8 * n)
            log(f"[bench] {name}: {ex[name]['median_ms']:.3f} ms, {ex[name]['GB/s']:.1f} GB/s")
        del x
        torch.cuda.empty_cache()

    if want("folds"):
        # ddot and dnrm2 at folds 2, 3, 4 (N = 2^28 wide-range), and dsum with incX = 2
        n = 1 << 28
        x = wide_gpu(n, 1, dev)
        y = wide_gpu(n, 2, dev)
        rec = {"ddot": {}, "dnrm2": {}}
        for fold in (2, 3, 4):
            rec["ddot"][str(fold)] = hbm_rec(time_each(
                lambda: L.reproblas_hip_rddot_device(fold, n, vp(x), 1, vp(y), 1, vp(out), sp), calls, WARM_HBM,
                stream), 16 * n)
            rec["dnrm2"][str(fold)] = hbm_rec(time_each(
                lambda: L.reproblas_hip_rdnrm2_device(fold, n, vp(x), 1, vp(out), sp), calls, WARM_HBM, stream),
                8 * n)
        h = n // 2
        r = hbm_rec(time_each(lambda: L.reproblas_hip_rdsum_device(3, h, vp(x), 2, vp(out), sp), calls, WARM_HBM,
                              stream), 8 * h)
        r.update({"terms": h, "GB/s_of_span": round(2 * r["GB/s"], 1)})
        rec["dsum_inc2"] = r
        ex["folds"] = ex["folds_2p28"] = rec
        log(f"[bench] folds: ddot {[rec['ddot'][f]['GB/s'] for f in rec['ddot']]}, "
            f"dnrm2 {[rec['dnrm2'][f]['GB/s'] for f in rec['dnrm2']]}")
        del x, y
        torch.cuda.empty_cache()

    if want("dgemv"):
        # dgemv 16384^2 column-major, fold 3, alpha 1, beta 0 and 0.5 (y normal(0,1))
        m = 16384
        A = torch.randn(m * m, generator=g, device=dev, dtype=torch.float64)
        xv = torch.randn(m, generator=g, device=dev, dtype=torch.float64)
        y0 = torch.randn(m, generator=g, device=dev, dtype=torch.float64)
        yv = y0.clone()
        for tr in ("N", "T"):
            for beta in (0.0, 0.5):
                fn = lambda: L.reproblas_hip_rdgemv_device(3, b"C", tr.encode(), m, m, 1.0, vp(A), m, vp(xv), 1,
                                                           beta, vp(yv), 1, sp)
                byts = 8 * (m * m + m + m * (2 if beta != 0.0 else 1))
                key = f"dgemv_{tr}_16384" + ("" if beta == 0.5 else "_beta0")
                ex[key] = hbm_rec(time_each(fn, calls, WARM_HBM, stream), byts)
                ex[key]["beta"] = beta
                log(f"[bench] dgemv {tr} beta {beta}: {ex[key]['median_ms']:.3f} ms, {ex[key]['GB/s']:.1f} GB/s")
        ex["dgemv"] = {"N": ex["dgemv_N_16384"], "T": ex["dgemv_T_16384"]}
        del A
        torch.cuda.empty_cache()

    if want("dgemm"):
        # dgemm 4096^3 column-major, fold 3, alpha 1: NN beta 0, NN beta 0.5 (C normal(0,1)), TN beta 0
        k = 4096
        A = torch.randn(k * k, generator=g, device=dev, dtype=torch.float64)
        B = torch.randn(k * k, generator=g, device=dev, dtype=torch.float64)
        C0 = torch.randn(k * k, generator=g, device=dev, dtype=torch.float64)
        C = C0.clone()
        ops = 11.0 * k ** 3  # (4*fold-1) ops per product, the reference's model (bench_rdgemm.c:125-127)
        for ta, beta, key in (("N", 0.0, "dgemm_NN_4096"), ("N", 0.5, "dgemm_NN_4096_beta05"),
                              ("T", 0.0, "dgemm_TN_4096")):
            fn = lambda: L.reproblas_hip_rdgemm_device(3, b"C", ta.encode(), b"N", k, k, k, 1.0, vp(A), k, vp(B), k,
                                                       beta, vp(C), k, sp)
            ex[key] = valu_rec(time_each(fn, 20, 2, stream), ops)
            ex[key]["products_per_s"] = round(k ** 3 / (ex[key]["median_ms"] / 1e3), 1)
            log(f"[bench] {key}: {ex[key]['median_ms']:.2f} ms, {ex[key]['valu_frac']:.3f} of the VALU issue rate")
        ex["dgemm"] = ex["dgemm_NN_4096"]
        del A, B, C, C0
        torch.cuda.empty_cache()
    if want("single"):
        ex["single"] = single(L, dev, stream, sp, vp, calls)
    if want("complex"):
        ex["complex"] = complex_path(L, dev, stream, sp, vp, calls)
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


def single(L, dev, stream, sp, vp, calls):
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
        rec[name + "_2p28"] = hbm_rec(time_each(fn, calls, 20, stream), byts)
        log(f"[bench] {name} 2^28: {rec[name + '_2p28']['median_ms']:.3f} ms, {rec[name + '_2p28']['GB/s']:.1f} GB/s")
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
        rec[f"sgemv_{tr}_16384"] = hbm_rec(time_each(fn, calls, 10, stream), 4 * (m * m + 3 * m))
        log(f"[bench] sgemv {tr}: {rec[f'sgemv_{tr}_16384']['median_ms']:.3f} ms")
    del A
    torch.cuda.empty_cache()
    k = 4096
    A = torch.randn(k * k, generator=g, device=dev, dtype=torch.float32)
    B = torch.randn(k * k, generator=g, device=dev, dtype=torch.float32)
    C = torch.zeros(k * k, device=dev, dtype=torch.float32)
    fn = lambda: L.reproblas_hip_rsgemm_device(3, b"C", b"N", b"N", k, k, k, 1.0, vp(A), k, vp(B), k, 0.0, vp(C), k,
                                               sp)
    r = valu_rec(time_each(fn, 20, 2, stream), 11.0 * k ** 3, VALU_PEAK_OPS_FP32)
    r.pop("frac_of_published_fp64")
    rec["sgemm_NN_4096"] = r
    log(f"[bench] sgemm: {r['median_ms']:.2f} ms, {r['valu_frac']:.3f} of the FP32 VALU issue rate")
    del A, B, C
    torch.cuda.empty_cache()
    import reproblas_amd as rb

    rb._check()
    return rec


def complex_path(L, dev, stream, sp, vp, calls):
    """Complex binned reductions at fold 3: zdotc, zsum, dznrm2 over N = 2^27
    double complex (wide-range parts), cdotc over 2^27 float complex, zgemv
    8192^2 column-major N and T (beta 0.5) and zgemm 2048^3 column-major NN."""
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
        rec[name + "_2p27"] = hbm_rec(time_each(fn, calls, 10, stream), byts)
        log(f"[bench] {name} 2^27: {rec[name + '_2p27']['median_ms']:.3f} ms")
    xf, yf = x.float(), y.float()
    del x, y
    torch.cuda.empty_cache()
    outf = torch.zeros(8, dtype=torch.float32, device=dev)
    rec["cdotc_2p27"] = hbm_rec(time_each(
        lambda: L.reproblas_hip_rcdotc_device(3, n, vp(xf), 1, vp(yf), 1, vp(outf), sp), calls, 20, stream), 16 * n)
    log(f"[bench] cdotc 2^27: {rec['cdotc_2p27']['median_ms']:.3f} ms")
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
        rec[f"zgemv_{tr}_8192"] = hbm_rec(time_each(fn, calls, 5, stream), 16 * (m * m + 3 * m))
        log(f"[bench] zgemv {tr} 8192: {rec[f'zgemv_{tr}_8192']['median_ms']:.3f} ms")
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
    rec["zgemm_NN_2048"] = valu_rec(time_each(fn, 20, 2, stream), 4 * 11.0 * k ** 3)  # four real products each
    log(f"[bench] zgemm 2048: {rec['zgemm_NN_2048']['median_ms']:.2f} ms")
    del A, B, C
    torch.cuda.empty_cache()
    import reproblas_amd as rb

    rb._check()
    return rec


def ragged_batch(L, dev, stream, sp, vp):
    """10,000 dot segments, lengths log-uniform in [16, 2^20] (CSR), wide-range
    values, plus 512 column-major NN gemms with M, N, K uniform in
    {32, 64, ..., 512}; fold 3. The autotuner picks the gemm tiling and the
    segment chunk; one graph launch of the plan is timed against the same
    work issued as per-call launches (each its own reproBLAS call) into
    separate outputs, and every output of the two is compared bit for bit."""
    import torch

    from reproblas_amd.autotune import tune_gemm_batch, tune_segments
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
    C2 = [torch.zeros_like(c) for _, _, c in triples]
    products = int(np.prod(sizes, axis=1).sum())
    log(f"[bench] ragged batch: {nseg} dot segments, {total} terms; 512 gemms, {products} products")
    tuned = tune_gemm_batch("N", "N", triples, use_cache=False, reps=3)
    tuned_seg = tune_segments("dot", off, x, y, use_cache=False, reps=3)
    b = Batch(3)
    b.add_segments("dot", off, x, y, res, chunk=tuned_seg["chunk"])
    b.add_gemms("N", "N", triples, tiling=tuned["tiling"])
    t_graph = time_each(lambda: b.launch(graph=True), 20, 2, stream)
    # the gemm branch alone: its VALU fraction inside the batch
    bg = Batch(3)
    bg.add_gemms("N", "N", triples, tiling=tuned["tiling"])
    t_gemm = time_each(lambda: bg.launch(graph=True), 10, 2, stream)
    bg.close()

    xp, yp, rp = x.data_ptr(), y.data_ptr(), res2.data_ptr()

    def per_call():
        for q in range(nseg):
            L.reproblas_hip_rddot_device(3, int(lens[q]), ctypes.c_void_p(xp + 8 * int(off[q])), 1,
                                         ctypes.c_void_p(yp + 8 * int(off[q])), 1, ctypes.c_void_p(rp + 8 * q), sp)
        for (M, N, K), (A, B, _), C in zip(sizes, triples, C2):
            L.reproblas_hip_rdgemm_device(3, b"C", b"N", b"N", int(M), int(N), int(K), 1.0, vp(A), int(M), vp(B),
                                          int(K), 0.0, vp(C), int(M), sp)

    t_calls = time_each(per_call, 3, 1, stream)
    torch.cuda.synchronize()
    same_seg = bool(torch.equal(res.view(torch.int64), res2.view(torch.int64)))
    same_gemm = all(bool(torch.equal(c.view(torch.int64), c2.view(torch.int64))) for (_, _, c), c2 in zip(triples, C2))
    b.close()
    tg, tc, tgm = median(t_graph), median(t_calls), median(t_gemm)
    log(f"[bench] ragged batch: graph {tg * 1e3:.2f} ms, per-call {tc * 1e3:.2f} ms, speedup {tc / tg:.2f}x, "
        f"identical segments={same_seg} gemms={same_gemm}; gemm branch alone {tgm * 1e3:.2f} ms")
    return {"graph": stats(t_graph), "per_call": stats(t_calls), "graph_ms": round(tg * 1e3, 3),
            "per_call_ms": round(tc * 1e3, 3), "speedup": round(tc / tg, 2), "segments": nseg,
            "segment_terms": total, "gemms": 512, "gemm_products": products, "tiling": tuned["tiling"],
            "tiling_ms": tuned["all_ms"], "chunk": tuned_seg["chunk"], "chunk_ms": tuned_seg["all_ms"],
            "gemm_branch_alone_ms": round(tgm * 1e3, 3),
            "gemm_branch_valu_frac": round(11.0 * products / tgm / VALU_PEAK_OPS, 4),
            "segments_identical": same_seg, "gemms_identical": same_gemm,
            "results_identical": same_seg and same_gemm}


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
    ts = time_each(lambda: ag.mlp_step(model, xb, tb), 10, 1, stream)
    t = median(ts)
    fwd = 2048 * (1024 * 4096 + 4096 * 4096 + 4096 * 1024)
    products = fwd + 2048 * (1024 * 4096 + 4096 * 4096 + 4096 * 1024) + 2048 * (4096 * 4096 + 4096 * 1024)
    log(f"[bench] mlp fwd+bwd: {t * 1e3:.1f} ms, {products / t / 1e12:.2f} T products/s")
    r = stats(ts)
    r.update({"ms": round(t * 1e3, 2), "products_per_step": products,
              "valu_frac": round(11.0 * products / t / VALU_PEAK_OPS, 4)})
    return r


if __name__ == "__main__":
    main()
