"""Autotuner for the batch and gemm launch shapes.

The knobs change only how work is cut (gemm output tile per lane: 4x4, 4x2 or
2x2; terms per segment chunk), never the bits: every candidate must produce
identical outputs, which the tuner checks before it trusts a timing. Results
are cached per device and workload signature in a JSON file
($REPROBLAS_AMD_TUNE_CACHE, default ~/.cache/reproblas_amd/autotune.json).

    best = tune_gemm_batch("N", "N", triples)          # {'tiling': 1, 'ms': ...}
    best = tune_segments("dot", offsets, x, y)         # {'chunk': 16384, ...}
    best = tune_gemm(M, N, K)                          # single-gemm tiling
"""
import hashlib
import json
import os

from .batch import Batch

CACHE = os.environ.get("REPROBLAS_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "reproblas_amd", "autotune.json"))


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    except OSError:
        pass


def _key(kind, sig):
    import torch

    dev = torch.cuda.get_device_name(0)
    return f"{dev}|{kind}|" + hashlib.sha1(json.dumps(sig, sort_keys=True).encode()).hexdigest()[:16]


def _time(fn, reps=3, warmup=1):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_gemm_batch(transa, transb, triples, fold=3, alpha=1.0, beta=0.0, candidates=(0, 1, 2), reps=3,
                    use_cache=True):
    """Pick the gemm tiling for a batch. C operands are restored after tuning."""
    import torch

    sig = {"t": transa + transb, "f": fold, "a": alpha, "b": beta,
           "shapes": [list(A.shape) + list(B.shape) for A, B, _ in triples]}
    key = _key("gemm_batch", sig)
    db = _load() if use_cache else {}
    if key in db:
        return db[key]
    saved = [C.clone() for _, _, C in triples]
    results, outs = {}, {}
    for t in candidates:
        for C, C0 in zip((c for _, _, c in triples), saved):
            C.copy_(C0)
        b = Batch(fold)
        b.add_gemms(transa, transb, triples, alpha, beta, tiling=t)
        b.launch(graph=True)
        torch.cuda.synchronize()
        outs[t] = [C.clone() for _, _, C in triples]
        results[t] = _time(lambda: b.launch(graph=True), reps)
        b.close()
    for C, C0 in zip((c for _, _, c in triples), saved):
        C.copy_(C0)
    ref = outs[candidates[0]]
    for t in candidates[1:]:
        for a, c in zip(ref, outs[t]):
            if not torch.equal(a.view(torch.int64), c.view(torch.int64)):
                raise RuntimeError(f"gemm tilings {candidates[0]} and {t} disagree: refusing to tune")
    best = min(results, key=results.get)
    rec = {"tiling": best, "ms": results[best], "all_ms": {str(k): v for k, v in results.items()}}
    if use_cache:
        db[key] = rec
        _store(db)
    return rec


def tune_segments(op, offsets, x, y=None, fold=3, candidates=(4096, 16384, 65536), reps=3, use_cache=True):
    """Pick the chunk size (terms per work item) for a segment batch."""
    import torch

    nseg = len(offsets) - 1
    sig = {"op": op, "f": fold, "n": int(offsets[-1]), "nseg": nseg,
           "h": hashlib.sha1(bytes(memoryview(__import__("numpy").asarray(offsets, dtype="int64")))).hexdigest()[:16]}
    key = _key("segments", sig)
    db = _load() if use_cache else {}
    if key in db:
        return db[key]
    results, outs = {}, {}
    for ch in candidates:
        res = torch.empty(nseg, dtype=torch.float64, device=x.device)
        b = Batch(fold)
        b.add_segments(op, offsets, x, y, res, chunk=ch)
        b.launch(graph=True)
        torch.cuda.synchronize()
        outs[ch] = res.clone()
        results[ch] = _time(lambda: b.launch(graph=True), reps)
        b.close()
    ref = outs[candidates[0]]
    for ch in candidates[1:]:
        if not torch.equal(ref.view(torch.int64), outs[ch].view(torch.int64)):
            raise RuntimeError(f"segment chunks {candidates[0]} and {ch} disagree: refusing to tune")
    best = min(results, key=results.get)
    rec = {"chunk": best, "ms": results[best], "all_ms": {str(k): v for k, v in results.items()}}
    if use_cache:
        db[key] = rec
        _store(db)
    return rec


def tune_gemm(M, N, K, fold=3, transa="N", transb="N", candidates=(0, 1, 2), reps=2, use_cache=True):
    """Pick the tiling for one gemm shape (column-major) and make it current."""
    import torch

    from . import rdgemm, set_gemm_tiling

    key = _key("gemm", {"M": M, "N": N, "K": K, "f": fold, "t": transa + transb})
    db = _load() if use_cache else {}
    if key not in db:
        g = torch.Generator(device="cuda")
        g.manual_seed(0)
        ar, ac = (M, K) if transa in "nN" else (K, M)
        br, bc = (K, N) if transb in "nN" else (N, K)
        A = torch.randn(ac, ar, generator=g, device="cuda", dtype=torch.float64)
        B = torch.randn(bc, br, generator=g, device="cuda", dtype=torch.float64)
        C = torch.zeros(N, M, device="cuda", dtype=torch.float64)
        results, outs = {}, {}
        for t in candidates:
            set_gemm_tiling(t)
            rdgemm(transa, transb, 1.0, A, B, 0.0, C, fold=fold)
            outs[t] = C.clone()
            results[t] = _time(lambda: rdgemm(transa, transb, 1.0, A, B, 0.0, C, fold=fold), reps)
        for t in candidates[1:]:
            if not torch.equal(outs[candidates[0]].view(torch.int64), outs[t].view(torch.int64)):
                raise RuntimeError(f"gemm tilings {candidates[0]} and {t} disagree: refusing to tune")
        best = min(results, key=results.get)
        db[key] = {"tiling": best, "ms": results[best], "all_ms": {str(k): v for k, v in results.items()}}
        if use_cache:
            _store(db)
    set_gemm_tiling(db[key]["tiling"])
    return db[key]
