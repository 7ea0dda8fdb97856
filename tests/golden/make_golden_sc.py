"""Writes tests/golden/sc.npz: single-precision and complex fixtures from the
compiled reference (oracle/_ref, built by `make -C oracle ref` where the
reference tree is present), so the single-precision and complex oracles stay
pinned where the reference library is absent.

    python tests/golden/make_golden_sc.py

Inputs are seeded draws (datagen.py); outputs are the reference library's
results and binned states. Data only.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import datagen as dg  # noqa: E402
from cpu_libs import reference_s, reference_z  # noqa: E402

NS = (1, 7, 257, 513, 1025)
FOLDS = (2, 3, 4)
KINDS = ("wide", "uniform", "big", "tiny", "mixed_range", "nan", "pinf", "both_inf")


def svec(kind, n, seed):
    if kind == "wide":
        return dg.wide_s(n, seed)
    if kind == "uniform":
        return dg.uniform(n, seed).astype(np.float32)
    if n < 2 and kind == "both_inf":
        kind = "nan"
    return dg.special_s(kind, n, seed)


def cvec(prec, kind, n, seed):
    dt = np.float64 if prec == "z" else np.float32
    if kind == "wide":
        v = dg.wide(2 * n, seed, -20, 20)
    elif kind == "uniform":
        v = dg.uniform(2 * n, seed)
    elif prec == "z":
        v = dg.special(kind, 2 * n, seed)
    else:
        v = svec(kind, 2 * n, seed)
    return np.ascontiguousarray(np.asarray(v, dtype=dt))


def main():
    R = reference_s("avx")
    if R is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree is present")
    out = {}
    # single precision BLAS1 and states
    for kind in KINDS:
        for n in NS:
            x, y = svec(kind, n, 100 + n), dg.wide_s(n, 200 + n)
            out[f"s_{kind}_{n}_x"], out[f"s_{kind}_{n}_y"] = x, y
            for fold in FOLDS:
                res = [R.rssum(fold, x), R.rsasum(fold, x), R.rsnrm2(fold, x), R.rsdot(fold, x, y)]
                out[f"s_{kind}_{n}_{fold}_res"] = np.array(res, np.float32)
                st = np.zeros(2 * fold, np.float32)
                R.sbssum(fold, x, st)
                out[f"s_{kind}_{n}_{fold}_sbssum"] = st
    # single precision gemv / gemm
    M, N, K = 37, 29, 300
    A = dg.wide_s(K * max(M, N) + 16, 1)
    B = dg.wide_s(K * max(M, N) + 16, 2)
    C0 = dg.normal(max(M, N) ** 2 + 16, 3).astype(np.float32)
    out["s_mat_A"], out["s_mat_B"], out["s_mat_C"] = A, B, C0
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    for order in "CR":
        for tr in "NT":
            y = C0.copy()
            lda = (M if order == "C" else N) + 1
            R.fn["rsgemv"](3, order.encode(), tr.encode(), M, N, 1.5, vp(A), lda, vp(B), 1, 0.5, vp(y), 1)
            out[f"s_gemv_{order}{tr}"] = y
            for tb in "NT":
                rm = order == "R"
                la = (K if tr == "N" else M) if rm else (M if tr == "N" else K)
                lb = (N if tb == "N" else K) if rm else (K if tb == "N" else N)
                c = C0.copy()
                R.fn["rsgemm"](3, order.encode(), tr.encode(), tb.encode(), M, N, K, 0.75, vp(A), la, vp(B), lb, 0.5,
                               vp(c), N if rm else M)
                out[f"s_gemm_{order}{tr}{tb}"] = c
    # complex
    for prec in "zc":
        Z = reference_z(prec, "avx")
        for kind in KINDS:
            for n in NS:
                x, y = cvec(prec, kind, n, 300 + n), cvec(prec, "wide", n, 400 + n)
                out[f"{prec}_{kind}_{n}_x"], out[f"{prec}_{kind}_{n}_y"] = x, y
                for fold in FOLDS:
                    res = np.concatenate([Z.rsum(fold, x), Z.rdot("u", fold, x, y), Z.rdot("c", fold, x, y),
                                          np.array([Z.rasum(fold, x), Z.rnrm2(fold, x)], x.dtype)])
                    out[f"{prec}_{kind}_{n}_{fold}_res"] = res
        Mz, Nz, Kz = 23, 17, 200
        dt = np.float64 if prec == "z" else np.float32
        Az = cvec(prec, "wide", Kz * max(Mz, Nz) + 8, 5)
        Bz = cvec(prec, "wide", Kz * max(Mz, Nz) + 8, 6)
        Cz = cvec(prec, "uniform", max(Mz, Nz) ** 2 + 8, 7)
        out[f"{prec}_mat_A"], out[f"{prec}_mat_B"], out[f"{prec}_mat_C"] = Az, Bz, Cz
        al, be = np.array([1.5, -0.5], dt), np.array([0.5, 0.25], dt)
        for order in "CR":
            for ta in "NTC":
                if not (order == "C" and ta == "C"):
                    y = Cz.copy()
                    lda = (Mz if order == "C" else Nz) + 1
                    Z.fn["rgemv"](3, order.encode(), ta.encode(), Mz, Nz, vp(al), vp(Az), lda, vp(Bz), 1, vp(be),
                                  vp(y), 1)
                    out[f"{prec}_gemv_{order}{ta}"] = y
                for tb in "NTC":
                    rm = order == "R"
                    la = (Kz if ta == "N" else Mz) if rm else (Mz if ta == "N" else Kz)
                    lb = (Nz if tb == "N" else Kz) if rm else (Kz if tb == "N" else Nz)
                    c = Cz.copy()
                    Z.fn["rgemm"](3, order.encode(), ta.encode(), tb.encode(), Mz, Nz, Kz, vp(al), vp(Az), la, vp(Bz),
                                  lb, vp(be), vp(c), Nz if rm else Mz)
                    out[f"{prec}_gemm_{order}{ta}{tb}"] = c
    path = os.path.join(HERE, "sc.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
