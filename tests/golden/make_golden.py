"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in a container that holds the reference tree: `make -C oracle ref` builds
oracle/_ref/libreproblas_ref.so from the reference's own C sources (AVX
sections, as its build selects on an AVX host), then

    python tests/golden/make_golden.py

stores seeded inputs and the reference library's outputs (results and binned
states) as .npz data. The fixtures carry data only; the GPU host needs
neither the reference tree nor its library to use them.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import datagen as dg  # noqa: E402
from cpu_libs import reference  # noqa: E402


def fills(n, seed):
    """The reference's test fills (tests/common/test_util.c:1413-1627) plus wide-range draws."""
    small, big = 2.0 ** -27, 2.0 ** 27
    out = {}
    out["uniform"] = dg.uniform(n, seed)
    out["wide"] = dg.wide(n, seed)
    out["normal"] = dg.normal(n, seed)
    out["mixed_range"] = dg.special("mixed_range", n, seed)
    out["big"] = dg.special("big", n, seed)
    out["tiny"] = dg.special("tiny", n, seed)
    out["constant"] = np.ones(n)
    v = np.full(n, small)
    v[0] = big
    out["pos_big"] = v
    v = np.full(n, small)
    v[0] = big
    v[-1] = -big if n > 1 else big
    out["pos_neg_big"] = v
    if n > 1:
        v = np.zeros(n)
        s = dg.rng(seed).uniform()
        for i in range(n // 2):
            v[i] = s
            v[n - 1 - i] = -s
            s = s * 2.0 ** (1 - 26)
        out["mountain"] = v
        h = (n + 1) // 2
        v = np.empty(n)
        v[:h] = np.sin(2.0 * np.pi * (np.arange(h) / n))
        v[0] = 0.0
        for i in range(h, n):
            v[i] = -v[i - n // 2]
        out["sine"] = v
    for k in ("nan", "pinf", "ninf", "both_inf", "inf_nan"):
        if n > 1 or k in ("nan", "pinf", "ninf"):
            out[k] = dg.special(k, n, seed)
    return out


def blas1(R):
    data = {}
    cases = []
    for n in (1, 7, 4095, 16411):
        fx = fills(n, 1000 + n)
        fy = dg.wide(n, 2000 + n)
        data[f"y_N{n}"] = fy
        for name, x in fx.items():
            if n == 16411 and name not in ("uniform", "wide", "big", "mixed_range"):
                continue
            key = f"{name}_N{n}"
            data["x_" + key] = x
            for fold in (2, 3, 4):
                for op in ("rdsum", "rdasum", "rdnrm2", "rddot"):
                    r = R.rddot(fold, x, fy) if op == "rddot" else getattr(R, op)(fold, x)
                    data[f"r_{op}_f{fold}_{key}"] = np.array([r])
                st = np.zeros(2 * fold)
                R.dbdsum(fold, x, st)
                data[f"b_dbdsum_f{fold}_{key}"] = st
                st = np.zeros(2 * fold)
                sc = R.dbdssq(fold, x, 0.0, st)
                data[f"b_dbdssq_f{fold}_{key}"] = st
                data[f"s_dbdssq_f{fold}_{key}"] = np.array([sc])
            cases.append(key)
    data["cases"] = np.array(cases)
    return data


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def level23(R):
    data = {}
    cases = []
    M, N = 257, 129
    for order in ("C", "R"):
        lda = (M if order == "C" else N) + 1
        if order == "C":
            A = dg.normal(34000, 301)
            data["A_gemv_C"] = A
        for trans in ("N", "T"):
            lx, ly = (N, M) if trans == "N" else (M, N)
            x = dg.normal(lx, 302)
            y0 = dg.normal(ly, 303)
            if order == "R":
                A = data["A_gemv_C"][: lda * M]
            data[f"x_gemv_{order}{trans}"] = x
            data[f"y0_gemv_{order}{trans}"] = y0
            for alpha, beta in ((1.0, 0.0), (1.0, 0.5), (2.0, 1.0), (0.0, 0.5)):
                for fold in (2, 3, 4):
                    y = y0.copy()
                    R.fn["rdgemv"](fold, order.encode(), trans.encode(), M, N, alpha, vp(A), lda, vp(x), 1, beta,
                                   vp(y), 1)
                    key = f"gemv_{order}{trans}_a{alpha}_b{beta}_f{fold}"
                    data["r_" + key] = y
                    cases.append(key)
    for (M, N, K) in ((64, 64, 64), (96, 80, 72)):
        big = max(M, N, K) ** 2
        data[f"A_gemm_{M}x{N}x{K}"] = dg.normal(big, 401 + M)
        data[f"B_gemm_{M}x{N}x{K}"] = dg.normal(big, 402 + M)
        data[f"C0_gemm_{M}x{N}x{K}"] = dg.normal(big, 403 + M)
        for order in ("C", "R"):
            for ta in ("N", "T"):
                for tb in ("N", "T"):
                    rm = order == "R"
                    ar, ac = (M, K) if ta == "N" else (K, M)
                    br, bc = (K, N) if tb == "N" else (N, K)
                    lda, ldb, ldc = (ac if rm else ar), (bc if rm else br), (N if rm else M)
                    A = data[f"A_gemm_{M}x{N}x{K}"][: lda * (ar if rm else ac)]
                    B = data[f"B_gemm_{M}x{N}x{K}"][: ldb * (br if rm else bc)]
                    C0 = data[f"C0_gemm_{M}x{N}x{K}"][: ldc * (M if rm else N)]
                    tag = f"gemm_{M}x{N}x{K}_{order}{ta}{tb}"
                    for alpha, beta in ((1.0, 0.0), (0.5, 1.0)) if M > 64 else ((1.0, 0.0), (1.0, 0.5), (0.5, 1.0)):
                        C = C0.copy()
                        R.fn["rdgemm"](3, order.encode(), ta.encode(), tb.encode(), M, N, K, alpha, vp(A), lda, vp(B),
                                       ldb, beta, vp(C), ldc)
                        key = f"{tag}_a{alpha}_b{beta}"
                        data["r_" + key] = C
                        cases.append(key)
    data["cases"] = np.array(cases)
    return data


def main():
    R = reference("avx")
    if R is None:
        sys.exit("oracle/_ref/libreproblas_ref.so is not built (make -C oracle ref, with the reference tree present)")
    np.savez_compressed(os.path.join(HERE, "blas1.npz"), **blas1(R))
    np.savez_compressed(os.path.join(HERE, "level23.npz"), **level23(R))
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
