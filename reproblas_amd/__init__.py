"""reproblas_amd -- reproducible (binned) FP64 and FP32 BLAS on AMD Instinct MI355X.

Python mirror of the reproBLAS C API (reference include/reproBLAS.h,
include/binnedBLAS.h) over the C ABI of ``libreproblas_hip.so``. Every
reduction runs in the library's gfx950 kernels; this module only marshals
arguments. Operands may be numpy arrays (staged to the GPU by the library) or
torch tensors on a ROCm device (used in place). There is no CPU compute path:
if the library is missing the import of any compute function fails loudly.

    import reproblas_amd as rb
    s = rb.rdsum(x)                    # fold 3, like reproBLAS_dsum
    d = rb.rddot(x, y, fold=2)         # like reproBLAS_rddot(2, ...)
    rb.rdgemm('N', 'N', 1.0, A, B, 0.0, C)   # column-major, in place on C
    t = rb.rssum(x32)                  # float32 input, float_binned state
"""
import ctypes
import os

__all__ = [
    "lib", "DEFAULT_FOLD", "ReproBLASError",
    "rdsum", "rdasum", "rdnrm2", "rddot", "rdgemv", "rdgemm",
    "dsum", "dasum", "dnrm2", "ddot", "dgemv", "dgemm",
    "dbdsum", "dbdasum", "dbddot", "dbdssq", "dbdgemv", "dbdgemm", "damax", "damaxm",
    "dbdconv", "ddbconv", "dbdbadd", "dbdbaddv", "dbsetzero", "dbnum",
    "gemm_tilings", "set_gemm_tiling", "get_gemm_tiling",
    "SDEFAULT_FOLD", "rssum", "rsasum", "rsnrm2", "rsdot", "rsgemv", "rsgemm",
    "ssum", "sasum", "snrm2", "sdot", "sgemv", "sgemm",
    "sbssum", "sbsasum", "sbsdot", "sbsssq", "sbsgemv", "sbsgemm", "samax", "samaxm",
    "sbsconv", "ssbconv", "sbsbadd", "sbsbaddv", "sbsetzero", "sbnum",
    "rzsum", "rzdotu", "rzdotc", "rdzasum", "rdznrm2", "rzgemv", "rzgemm",
    "rcsum", "rcdotu", "rcdotc", "rscasum", "rscnrm2", "rcgemv", "rcgemm",
    "zbzsum", "zbzdotu", "zbzdotc", "dbzasum", "dbzssq", "cbcsum", "cbcdotu", "cbcdotc", "sbcasum", "sbcssq",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libreproblas_hip.so")
DEFAULT_FOLD = 3  # reference config.h:14 (DIDEFAULTFOLD)
SDEFAULT_FOLD = 3  # reference config.h:22 (SIDEFAULTFOLD)


class ReproBLASError(RuntimeError):
    pass


_lib = None
_c_int, _c_i64, _c_double, _c_char, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_char, ctypes.c_void_p
_c_float = ctypes.c_float

_SIGS = {
    # reproBLAS.h
    "reproBLAS_rdsum": (_c_double, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rdasum": (_c_double, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rdnrm2": (_c_double, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rddot": (_c_double, [_c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    "reproBLAS_rdgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_double, _vp, _c_int, _vp, _c_int,
                                _c_double, _vp, _c_int]),
    "reproBLAS_rdgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_double, _vp, _c_int,
                                _vp, _c_int, _c_double, _vp, _c_int]),
    "reproBLAS_dsum": (_c_double, [_c_int, _vp, _c_int]),
    "reproBLAS_dasum": (_c_double, [_c_int, _vp, _c_int]),
    "reproBLAS_dnrm2": (_c_double, [_c_int, _vp, _c_int]),
    "reproBLAS_ddot": (_c_double, [_c_int, _vp, _c_int, _vp, _c_int]),
    "reproBLAS_dgemv": (None, [_c_char, _c_char, _c_int, _c_int, _c_double, _vp, _c_int, _vp, _c_int, _c_double,
                               _vp, _c_int]),
    "reproBLAS_dgemm": (None, [_c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_double, _vp, _c_int, _vp,
                               _c_int, _c_double, _vp, _c_int]),
    # binnedBLAS.h
    "binnedBLAS_damax": (_c_double, [_c_int, _vp, _c_int]),
    "binnedBLAS_damaxm": (_c_double, [_c_int, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_dbdsum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_dbdasum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_dbddot": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_dbdssq": (_c_double, [_c_int, _c_int, _vp, _c_int, _c_double, _vp]),
    "binnedBLAS_dmdsum": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_dmdasum": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_dmddot": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_dmdssq": (_c_double, [_c_int, _c_int, _vp, _c_int, _c_double, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_dbdgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_double, _vp, _c_int, _vp, _c_int,
                                  _vp, _c_int]),
    "binnedBLAS_dbdgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_double, _vp, _c_int,
                                  _vp, _c_int, _vp, _c_int]),
    # binned.h (subset used from Python)
    "binned_dbnum": (_c_int, [_c_int]),
    "binned_dbsetzero": (None, [_c_int, _vp]),
    "binned_dbdconv": (None, [_c_int, _c_double, _vp]),
    "binned_ddbconv": (_c_double, [_c_int, _vp]),
    "binned_dbdbadd": (None, [_c_int, _vp, _vp]),
    "binned_dbdbaddv": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    # reproblas_hip.h
    "reproblas_hip_last_error": (_c_int, []),
    "reproblas_hip_last_error_string": (ctypes.c_char_p, []),
    "reproblas_hip_clear_error": (None, []),
    "reproblas_hip_set_stream": (None, [_vp]),
    "reproblas_hip_get_stream": (_vp, []),
    "reproblas_hip_min_fold": (_c_int, []),
    "reproblas_hip_max_fold": (_c_int, []),
    "reproblas_hip_rdsum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rdasum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rdnrm2_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rddot_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_dbdsum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_dbdasum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_dbddot_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_dbdssq_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp, _vp]),
    "reproblas_hip_rdgemv_device": (_c_int, [_c_int, _c_char, _c_char, _c_i64, _c_i64, _c_double, _vp, _c_i64, _vp,
                                             _c_i64, _c_double, _vp, _c_i64, _vp]),
    "reproblas_hip_rdgemm_device": (_c_int, [_c_int, _c_char, _c_char, _c_char, _c_i64, _c_i64, _c_i64, _c_double,
                                             _vp, _c_i64, _vp, _c_i64, _c_double, _vp, _c_i64, _vp]),
    # single precision (reproBLAS.h, binnedBLAS.h, binned.h, reproblas_hip.h)
    "reproBLAS_rssum": (_c_float, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rsasum": (_c_float, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rsnrm2": (_c_float, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rsdot": (_c_float, [_c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    "reproBLAS_rsgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_float, _vp, _c_int, _vp, _c_int,
                                _c_float, _vp, _c_int]),
    "reproBLAS_rsgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_float, _vp, _c_int,
                                _vp, _c_int, _c_float, _vp, _c_int]),
    "binnedBLAS_samax": (_c_float, [_c_int, _vp, _c_int]),
    "binnedBLAS_samaxm": (_c_float, [_c_int, _vp, _c_int, _vp, _c_int]),
    "binnedBLAS_sbssum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_sbsasum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_sbsdot": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_sbsssq": (_c_float, [_c_int, _c_int, _vp, _c_int, _c_float, _vp]),
    "binnedBLAS_sbsgemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _c_float, _vp, _c_int, _vp, _c_int,
                                  _vp, _c_int]),
    "binnedBLAS_sbsgemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _c_float, _vp, _c_int,
                                  _vp, _c_int, _vp, _c_int]),
    "binned_sbnum": (_c_int, [_c_int]),
    "binned_sbsetzero": (None, [_c_int, _vp]),
    "binned_sbsconv": (None, [_c_int, _c_float, _vp]),
    "binned_ssbconv": (_c_float, [_c_int, _vp]),
    "binned_sbsbadd": (None, [_c_int, _vp, _vp]),
    "binned_sbsbaddv": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    "reproblas_hip_rssum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rsasum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rsnrm2_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_rsdot_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_sbssum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_sbsasum_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_sbsdot_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp]),
    "reproblas_hip_sbsssq_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp, _vp]),
    "reproblas_hip_rsgemv_device": (_c_int, [_c_int, _c_char, _c_char, _c_i64, _c_i64, _c_float, _vp, _c_i64, _vp,
                                             _c_i64, _c_float, _vp, _c_i64, _vp]),
    "reproblas_hip_rsgemm_device": (_c_int, [_c_int, _c_char, _c_char, _c_char, _c_i64, _c_i64, _c_i64, _c_float,
                                             _vp, _c_i64, _vp, _c_i64, _c_float, _vp, _c_i64, _vp]),
    # complex (reproBLAS.h, binnedBLAS.h, reproblas_hip.h); pointers to interleaved pairs
    **{f"reproBLAS_r{z}sum_sub": (None, [_c_int, _c_int, _vp, _c_int, _vp]) for z in "zc"},
    **{f"reproBLAS_r{z}dot{k}_sub": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp]) for z in "zc" for k in "uc"},
    "reproBLAS_rdzasum": (_c_double, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rdznrm2": (_c_double, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rscasum": (_c_float, [_c_int, _c_int, _vp, _c_int]),
    "reproBLAS_rscnrm2": (_c_float, [_c_int, _c_int, _vp, _c_int]),
    **{f"reproBLAS_r{z}gemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _vp,
                                      _vp, _c_int]) for z in "zc"},
    **{f"reproBLAS_r{z}gemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _vp, _vp, _c_int,
                                      _vp, _c_int, _vp, _vp, _c_int]) for z in "zc"},
    **{f"binnedBLAS_{z}b{z}sum": (None, [_c_int, _c_int, _vp, _c_int, _vp]) for z in "zc"},
    **{f"binnedBLAS_{z}b{z}dot{k}": (None, [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp]) for z in "zc" for k in "uc"},
    "binnedBLAS_dbzasum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_sbcasum": (None, [_c_int, _c_int, _vp, _c_int, _vp]),
    "binnedBLAS_dbzssq": (_c_double, [_c_int, _c_int, _vp, _c_int, _c_double, _vp]),
    "binnedBLAS_sbcssq": (_c_float, [_c_int, _c_int, _vp, _c_int, _c_float, _vp]),
    **{f"binnedBLAS_{z}b{z}gemv": (None, [_c_int, _c_char, _c_char, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int,
                                          _vp, _c_int]) for z in "zc"},
    **{f"binnedBLAS_{z}b{z}gemm": (None, [_c_int, _c_char, _c_char, _c_char, _c_int, _c_int, _c_int, _vp, _vp,
                                          _c_int, _vp, _c_int, _vp, _c_int]) for z in "zc"},
    **{f"reproblas_hip_r{z}{op}_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _vp])
       for z, op in (("z", "sum"), ("dz", "asum"), ("dz", "nrm2"), ("c", "sum"), ("sc", "asum"), ("sc", "nrm2"))},
    **{f"reproblas_hip_r{z}dot{k}_device": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp])
       for z in "zc" for k in "uc"},
    **{f"reproblas_hip_r{z}gemv_device": (_c_int, [_c_int, _c_char, _c_char, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp,
                                                   _c_i64, _vp, _vp, _c_i64, _vp]) for z in "zc"},
    **{f"reproblas_hip_r{z}gemm_device": (_c_int, [_c_int, _c_char, _c_char, _c_char, _c_i64, _c_i64, _c_i64, _vp,
                                                   _vp, _c_i64, _vp, _c_i64, _vp, _vp, _c_i64, _vp]) for z in "zc"},
    "reproblas_hip_gemm_tiling_count": (_c_int, []),
    "reproblas_hip_gemm_set_tiling": (_c_int, [_c_int]),
    "reproblas_hip_gemm_get_tiling": (_c_int, []),
}


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ReproBLASError(
                f"{LIB_PATH} is not built: run `python -m reproblas_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ------------------------------------------------------------- marshalling --
def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _ptr(a, dtype="float64"):
    """Address of the first element of a numpy array or torch tensor of ``dtype``."""
    if a is None:
        return None
    if _is_torch(a):
        if str(a.dtype) != "torch." + dtype:
            raise TypeError(f"this call works on {dtype} tensors, got {a.dtype}")
        if not a.is_contiguous():
            raise ValueError("reproblas_amd needs contiguous tensors (pass strides through inc/ld arguments)")
        return ctypes.c_void_p(a.data_ptr())
    import numpy as np
    if not isinstance(a, np.ndarray) or a.dtype != np.dtype(dtype):
        raise TypeError(f"this call works on {dtype} numpy arrays or torch tensors")
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("reproblas_amd needs contiguous arrays (pass strides through inc/ld arguments)")
    return ctypes.c_void_p(a.ctypes.data)


def _bind_stream(*arrays):
    """Route the synchronous calls onto torch's current stream when operands are torch tensors."""
    for a in arrays:
        if a is not None and _is_torch(a) and a.is_cuda:
            import torch
            lib().reproblas_hip_set_stream(ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream))
            return
    lib().reproblas_hip_set_stream(None)


def _check():
    L = lib()
    e = L.reproblas_hip_last_error()
    if e != 0:
        msg = L.reproblas_hip_last_error_string().decode()
        L.reproblas_hip_clear_error()
        raise ReproBLASError(f"reproblas_hip error {e}: {msg}")


def _numel(a):
    return a.numel() if _is_torch(a) else a.size


def _n_of(a, inc):
    inc = abs(inc) or 1
    return (_numel(a) + inc - 1) // inc


# --------------------------------------------------------------------- BLAS1 --
def rdsum(x, fold=DEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rdsum: reproducible sum of x[0], x[inc], ... (n terms)."""
    _bind_stream(x)
    r = lib().reproBLAS_rdsum(fold, _n_of(x, inc) if n is None else n, _ptr(x), inc)
    _check()
    return r


def rdasum(x, fold=DEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rdasum: reproducible sum of |x_i|."""
    _bind_stream(x)
    r = lib().reproBLAS_rdasum(fold, _n_of(x, inc) if n is None else n, _ptr(x), inc)
    _check()
    return r


def rdnrm2(x, fold=DEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rdnrm2: reproducible Euclidean norm."""
    _bind_stream(x)
    r = lib().reproBLAS_rdnrm2(fold, _n_of(x, inc) if n is None else n, _ptr(x), inc)
    _check()
    return r


def rddot(x, y, fold=DEFAULT_FOLD, incx=1, incy=1, n=None):
    """reproBLAS_rddot: reproducible sum of fl(x_i*y_i)."""
    _bind_stream(x, y)
    if n is None:
        n = min(_n_of(x, incx), _n_of(y, incy))
    r = lib().reproBLAS_rddot(fold, n, _ptr(x), incx, _ptr(y), incy)
    _check()
    return r


def dsum(x, inc=1):
    return rdsum(x, DEFAULT_FOLD, inc)


def dasum(x, inc=1):
    return rdasum(x, DEFAULT_FOLD, inc)


def dnrm2(x, inc=1):
    return rdnrm2(x, DEFAULT_FOLD, inc)


def ddot(x, y, incx=1, incy=1):
    return rddot(x, y, DEFAULT_FOLD, incx, incy)


# ------------------------------------------------------- binned-state BLAS1 --
def dbnum(fold=DEFAULT_FOLD):
    return 2 * fold


def dbdsum(Y, x, fold=DEFAULT_FOLD, inc=1):
    """binnedBLAS_dbdsum: Y (a 2*fold float64 binned value) += sum(x)."""
    _bind_stream(x, Y)
    lib().binnedBLAS_dbdsum(fold, _n_of(x, inc), _ptr(x), inc, _ptr(Y))
    _check()


def dbdasum(Y, x, fold=DEFAULT_FOLD, inc=1):
    _bind_stream(x, Y)
    lib().binnedBLAS_dbdasum(fold, _n_of(x, inc), _ptr(x), inc, _ptr(Y))
    _check()


def dbddot(Z, x, y, fold=DEFAULT_FOLD, incx=1, incy=1):
    _bind_stream(x, y, Z)
    n = min(_n_of(x, incx), _n_of(y, incy))
    lib().binnedBLAS_dbddot(fold, n, _ptr(x), incx, _ptr(y), incy, _ptr(Z))
    _check()


def dbdssq(Y, x, scale=0.0, fold=DEFAULT_FOLD, inc=1):
    """binnedBLAS_dbdssq: returns the new scale; Y accumulates the scaled squares."""
    _bind_stream(x, Y)
    r = lib().binnedBLAS_dbdssq(fold, _n_of(x, inc), _ptr(x), inc, scale, _ptr(Y))
    _check()
    return r


def damax(x, inc=1):
    _bind_stream(x)
    r = lib().binnedBLAS_damax(_n_of(x, inc), _ptr(x), inc)
    _check()
    return r


def damaxm(x, y, incx=1, incy=1):
    _bind_stream(x, y)
    r = lib().binnedBLAS_damaxm(min(_n_of(x, incx), _n_of(y, incy)), _ptr(x), incx, _ptr(y), incy)
    _check()
    return r


# --------------------------------------------------------- binned helpers --
def dbsetzero(Y, fold=DEFAULT_FOLD):
    lib().binned_dbsetzero(fold, _ptr(Y))


def dbdconv(x, fold=DEFAULT_FOLD):
    import numpy as np
    Y = np.zeros(2 * fold)
    lib().binned_dbdconv(fold, x, _ptr(Y))
    return Y


def ddbconv(Y, fold=DEFAULT_FOLD):
    return lib().binned_ddbconv(fold, _ptr(Y))


def dbdbadd(X, Y, fold=DEFAULT_FOLD):
    lib().binned_dbdbadd(fold, _ptr(X), _ptr(Y))


def dbdbaddv(X, Y, n, fold=DEFAULT_FOLD, incx=1, incy=1):
    _bind_stream(X, Y)
    lib().binned_dbdbaddv(fold, n, _ptr(X), incx, _ptr(Y), incy)
    _check()


# ------------------------------------------------------------------ level 2/3 --
def _ch(c):
    return c.encode() if isinstance(c, str) else c


def rdgemv(trans, alpha, A, x, beta, y, fold=DEFAULT_FOLD, order="C", M=None, N=None, lda=None, incx=1, incy=1):
    """reproBLAS_rdgemv on a 2-D A; y is updated in place.

    ``order`` 'C' means column-major (the reference's default for anything but
    'r'/'R'), so a C-contiguous 2-D array of shape (n, m) is an m x n
    column-major matrix. M, N, lda default from that view."""
    rm = order in ("r", "R")
    if M is None or N is None or lda is None:
        sh = tuple(A.shape)
        if rm:
            M, N, lda = sh[0], sh[1], sh[1]
        else:
            M, N, lda = sh[1], sh[0], sh[1]
    _bind_stream(A, x, y)
    lib().reproBLAS_rdgemv(fold, _ch(order), _ch(trans), M, N, alpha, _ptr(A), lda, _ptr(x), incx, beta, _ptr(y),
                           incy)
    _check()


def rdgemm(transa, transb, alpha, A, B, beta, C, fold=DEFAULT_FOLD, order="C", M=None, N=None, K=None,
           lda=None, ldb=None, ldc=None):
    """reproBLAS_rdgemm; C is updated in place. Shapes as for :func:`rdgemv`."""
    rm = order in ("r", "R")

    def dims(X):
        s = tuple(X.shape)
        return (s[0], s[1], s[1]) if rm else (s[1], s[0], s[1])

    ra, ca, la = dims(A)
    rb_, cb, lb = dims(B)
    rc, cc, lc = dims(C)
    na = transa in ("n", "N")
    M = (ra if na else ca) if M is None else M
    K = (ca if na else ra) if K is None else K
    N = (cb if transb in ("n", "N") else rb_) if N is None else N
    lda, ldb, ldc = lda or la, ldb or lb, ldc or lc
    _bind_stream(A, B, C)
    lib().reproBLAS_rdgemm(fold, _ch(order), _ch(transa), _ch(transb), M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb,
                           beta, _ptr(C), ldc)
    _check()


def dgemv(trans, alpha, A, x, beta, y, **kw):
    return rdgemv(trans, alpha, A, x, beta, y, fold=DEFAULT_FOLD, **kw)


def dgemm(transa, transb, alpha, A, B, beta, C, **kw):
    return rdgemm(transa, transb, alpha, A, B, beta, C, fold=DEFAULT_FOLD, **kw)


def dbdgemv(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incy=1):
    _bind_stream(A, x, Yb)
    lib().binnedBLAS_dbdgemv(fold, _ch(order), _ch(trans), M, N, alpha, _ptr(A), lda, _ptr(x), incx, _ptr(Yb), incy)
    _check()


def dbdgemm(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, Cb, ldc):
    _bind_stream(A, B, Cb)
    lib().binnedBLAS_dbdgemm(fold, _ch(order), _ch(transa), _ch(transb), M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb,
                             _ptr(Cb), ldc)
    _check()


def gemm_tilings():
    return lib().reproblas_hip_gemm_tiling_count()


def set_gemm_tiling(t):
    if lib().reproblas_hip_gemm_set_tiling(t) != 0:
        _check()


def get_gemm_tiling():
    return lib().reproblas_hip_gemm_get_tiling()


# ===================================================== single precision --
# float32 inputs, float_binned states of 2*fold float32 (folds 2..21);
# reference include/reproBLAS.h:32-35,61,66 and binnedBLAS.h:37-104.
def _sp(a):
    return _ptr(a, "float32")


def rssum(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rssum: reproducible float sum."""
    _bind_stream(x)
    r = lib().reproBLAS_rssum(fold, _n_of(x, inc) if n is None else n, _sp(x), inc)
    _check()
    return r


def rsasum(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    _bind_stream(x)
    r = lib().reproBLAS_rsasum(fold, _n_of(x, inc) if n is None else n, _sp(x), inc)
    _check()
    return r


def rsnrm2(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    _bind_stream(x)
    r = lib().reproBLAS_rsnrm2(fold, _n_of(x, inc) if n is None else n, _sp(x), inc)
    _check()
    return r


def rsdot(x, y, fold=SDEFAULT_FOLD, incx=1, incy=1, n=None):
    _bind_stream(x, y)
    if n is None:
        n = min(_n_of(x, incx), _n_of(y, incy))
    r = lib().reproBLAS_rsdot(fold, n, _sp(x), incx, _sp(y), incy)
    _check()
    return r


def ssum(x, inc=1):
    return rssum(x, SDEFAULT_FOLD, inc)


def sasum(x, inc=1):
    return rsasum(x, SDEFAULT_FOLD, inc)


def snrm2(x, inc=1):
    return rsnrm2(x, SDEFAULT_FOLD, inc)


def sdot(x, y, incx=1, incy=1):
    return rsdot(x, y, SDEFAULT_FOLD, incx, incy)


def sbnum(fold=SDEFAULT_FOLD):
    return 2 * fold


def sbssum(Y, x, fold=SDEFAULT_FOLD, inc=1):
    """binnedBLAS_sbssum: Y (a 2*fold float32 binned value) += sum(x)."""
    _bind_stream(x, Y)
    lib().binnedBLAS_sbssum(fold, _n_of(x, inc), _sp(x), inc, _sp(Y))
    _check()


def sbsasum(Y, x, fold=SDEFAULT_FOLD, inc=1):
    _bind_stream(x, Y)
    lib().binnedBLAS_sbsasum(fold, _n_of(x, inc), _sp(x), inc, _sp(Y))
    _check()


def sbsdot(Z, x, y, fold=SDEFAULT_FOLD, incx=1, incy=1):
    _bind_stream(x, y, Z)
    n = min(_n_of(x, incx), _n_of(y, incy))
    lib().binnedBLAS_sbsdot(fold, n, _sp(x), incx, _sp(y), incy, _sp(Z))
    _check()


def sbsssq(Y, x, scale=0.0, fold=SDEFAULT_FOLD, inc=1):
    """binnedBLAS_sbsssq: returns the new scale; Y accumulates the scaled squares."""
    _bind_stream(x, Y)
    r = lib().binnedBLAS_sbsssq(fold, _n_of(x, inc), _sp(x), inc, scale, _sp(Y))
    _check()
    return r


def samax(x, inc=1):
    _bind_stream(x)
    r = lib().binnedBLAS_samax(_n_of(x, inc), _sp(x), inc)
    _check()
    return r


def samaxm(x, y, incx=1, incy=1):
    _bind_stream(x, y)
    r = lib().binnedBLAS_samaxm(min(_n_of(x, incx), _n_of(y, incy)), _sp(x), incx, _sp(y), incy)
    _check()
    return r


def sbsetzero(Y, fold=SDEFAULT_FOLD):
    lib().binned_sbsetzero(fold, _sp(Y))


def sbsconv(x, fold=SDEFAULT_FOLD):
    import numpy as np
    Y = np.zeros(2 * fold, np.float32)
    lib().binned_sbsconv(fold, x, _sp(Y))
    return Y


def ssbconv(Y, fold=SDEFAULT_FOLD):
    return lib().binned_ssbconv(fold, _sp(Y))


def sbsbadd(X, Y, fold=SDEFAULT_FOLD):
    lib().binned_sbsbadd(fold, _sp(X), _sp(Y))


def sbsbaddv(X, Y, n, fold=SDEFAULT_FOLD, incx=1, incy=1):
    _bind_stream(X, Y)
    lib().binned_sbsbaddv(fold, n, _sp(X), incx, _sp(Y), incy)
    _check()


def rsgemv(trans, alpha, A, x, beta, y, fold=SDEFAULT_FOLD, order="C", M=None, N=None, lda=None, incx=1, incy=1):
    """reproBLAS_rsgemv on a 2-D float32 A (shapes as for :func:`rdgemv`)."""
    rm = order in ("r", "R")
    if M is None or N is None or lda is None:
        sh = tuple(A.shape)
        if rm:
            M, N, lda = sh[0], sh[1], sh[1]
        else:
            M, N, lda = sh[1], sh[0], sh[1]
    _bind_stream(A, x, y)
    lib().reproBLAS_rsgemv(fold, _ch(order), _ch(trans), M, N, alpha, _sp(A), lda, _sp(x), incx, beta, _sp(y),
                           incy)
    _check()


def rsgemm(transa, transb, alpha, A, B, beta, C, fold=SDEFAULT_FOLD, order="C", M=None, N=None, K=None,
           lda=None, ldb=None, ldc=None):
    """reproBLAS_rsgemm on float32 operands; C is updated in place."""
    rm = order in ("r", "R")

    def dims(X):
        s = tuple(X.shape)
        return (s[0], s[1], s[1]) if rm else (s[1], s[0], s[1])

    ra, ca, la = dims(A)
    rb_, cb, lb = dims(B)
    rc, cc, lc = dims(C)
    na = transa in ("n", "N")
    M = (ra if na else ca) if M is None else M
    K = (ca if na else ra) if K is None else K
    N = (cb if transb in ("n", "N") else rb_) if N is None else N
    lda, ldb, ldc = lda or la, ldb or lb, ldc or lc
    _bind_stream(A, B, C)
    lib().reproBLAS_rsgemm(fold, _ch(order), _ch(transa), _ch(transb), M, N, K, alpha, _sp(A), lda, _sp(B), ldb,
                           beta, _sp(C), ldc)
    _check()


def sgemv(trans, alpha, A, x, beta, y, **kw):
    return rsgemv(trans, alpha, A, x, beta, y, fold=SDEFAULT_FOLD, **kw)


def sgemm(transa, transb, alpha, A, B, beta, C, **kw):
    return rsgemm(transa, transb, alpha, A, B, beta, C, fold=SDEFAULT_FOLD, **kw)


def sbsgemv(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incy=1):
    _bind_stream(A, x, Yb)
    lib().binnedBLAS_sbsgemv(fold, _ch(order), _ch(trans), M, N, alpha, _sp(A), lda, _sp(x), incx, _sp(Yb), incy)
    _check()


def sbsgemm(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, Cb, ldc):
    _bind_stream(A, B, Cb)
    lib().binnedBLAS_sbsgemm(fold, _ch(order), _ch(transa), _ch(transb), M, N, K, alpha, _sp(A), lda, _sp(B), ldb,
                             _sp(Cb), ldc)
    _check()


# ============================================================ complex --
# complex128 ("z", double_complex_binned) and complex64 ("c",
# float_complex_binned) operands; increments count complex elements.
# Reference include/reproBLAS.h:37-47,73-90 and binnedBLAS.h:39-128.
def _zp(a, prec):
    return _ptr(a, "complex128" if prec == "z" else "complex64")


def _zres(prec):
    import numpy as np
    return np.zeros(1, np.complex128 if prec == "z" else np.complex64)


def _zpair(v, prec):
    import numpy as np
    return np.array([v], dtype=np.complex128 if prec == "z" else np.complex64)


def _rzsum(prec, x, fold, inc, n):
    _bind_stream(x)
    out = _zres(prec)
    getattr(lib(), f"reproBLAS_r{prec}sum_sub")(fold, _n_of(x, inc) if n is None else n, _zp(x, prec), inc,
                                                _ptr(out, out.dtype.name))
    _check()
    return complex(out[0])


def _rzdot(prec, kind, x, y, fold, incx, incy, n):
    _bind_stream(x, y)
    if n is None:
        n = min(_n_of(x, incx), _n_of(y, incy))
    out = _zres(prec)
    getattr(lib(), f"reproBLAS_r{prec}dot{kind}_sub")(fold, n, _zp(x, prec), incx, _zp(y, prec), incy,
                                                      _ptr(out, out.dtype.name))
    _check()
    return complex(out[0])


def rzsum(x, fold=DEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rzsum_sub: reproducible complex sum (each component binned)."""
    return _rzsum("z", x, fold, inc, n)


def rzdotu(x, y, fold=DEFAULT_FOLD, incx=1, incy=1, n=None):
    """reproBLAS_rzdotu_sub: sum of x_i*y_i."""
    return _rzdot("z", "u", x, y, fold, incx, incy, n)


def rzdotc(x, y, fold=DEFAULT_FOLD, incx=1, incy=1, n=None):
    """reproBLAS_rzdotc_sub: sum of conj(x_i)*y_i."""
    return _rzdot("z", "c", x, y, fold, incx, incy, n)


def rdzasum(x, fold=DEFAULT_FOLD, inc=1, n=None):
    """reproBLAS_rdzasum: sum of |re| + |im|."""
    _bind_stream(x)
    r = lib().reproBLAS_rdzasum(fold, _n_of(x, inc) if n is None else n, _zp(x, "z"), inc)
    _check()
    return r


def rdznrm2(x, fold=DEFAULT_FOLD, inc=1, n=None):
    _bind_stream(x)
    r = lib().reproBLAS_rdznrm2(fold, _n_of(x, inc) if n is None else n, _zp(x, "z"), inc)
    _check()
    return r


def rcsum(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    return _rzsum("c", x, fold, inc, n)


def rcdotu(x, y, fold=SDEFAULT_FOLD, incx=1, incy=1, n=None):
    return _rzdot("c", "u", x, y, fold, incx, incy, n)


def rcdotc(x, y, fold=SDEFAULT_FOLD, incx=1, incy=1, n=None):
    return _rzdot("c", "c", x, y, fold, incx, incy, n)


def rscasum(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    _bind_stream(x)
    r = lib().reproBLAS_rscasum(fold, _n_of(x, inc) if n is None else n, _zp(x, "c"), inc)
    _check()
    return r


def rscnrm2(x, fold=SDEFAULT_FOLD, inc=1, n=None):
    _bind_stream(x)
    r = lib().reproBLAS_rscnrm2(fold, _n_of(x, inc) if n is None else n, _zp(x, "c"), inc)
    _check()
    return r


def _zstate(prec, kind, Y, x, y, fold, inc):
    _bind_stream(x, y, Y)
    rdt = "float64" if prec == "z" else "float32"
    L = lib()
    if kind == "sum":
        getattr(L, f"binnedBLAS_{prec}b{prec}sum")(fold, _n_of(x, inc), _zp(x, prec), inc, _ptr(Y, rdt))
    elif kind in ("u", "c"):
        n = min(_n_of(x, inc), _n_of(y, inc))
        getattr(L, f"binnedBLAS_{prec}b{prec}dot{kind}")(fold, n, _zp(x, prec), inc, _zp(y, prec), inc, _ptr(Y, rdt))
    else:
        getattr(L, ("binnedBLAS_dbzasum" if prec == "z" else "binnedBLAS_sbcasum"))(fold, _n_of(x, inc),
                                                                                   _zp(x, prec), inc, _ptr(Y, rdt))
    _check()


def zbzsum(Y, x, fold=DEFAULT_FOLD, inc=1):
    """binnedBLAS_zbzsum: Y (4*fold float64 complex binned value) += sum(x)."""
    _zstate("z", "sum", Y, x, None, fold, inc)


def zbzdotu(Z, x, y, fold=DEFAULT_FOLD, inc=1):
    _zstate("z", "u", Z, x, y, fold, inc)


def zbzdotc(Z, x, y, fold=DEFAULT_FOLD, inc=1):
    _zstate("z", "c", Z, x, y, fold, inc)


def dbzasum(Y, x, fold=DEFAULT_FOLD, inc=1):
    """binnedBLAS_dbzasum: Y (2*fold float64 binned value) += sum(|re| + |im|)."""
    _zstate("z", "asum", Y, x, None, fold, inc)


def dbzssq(Y, x, scale=0.0, fold=DEFAULT_FOLD, inc=1):
    _bind_stream(x, Y)
    r = lib().binnedBLAS_dbzssq(fold, _n_of(x, inc), _zp(x, "z"), inc, scale, _ptr(Y))
    _check()
    return r


def cbcsum(Y, x, fold=SDEFAULT_FOLD, inc=1):
    _zstate("c", "sum", Y, x, None, fold, inc)


def cbcdotu(Z, x, y, fold=SDEFAULT_FOLD, inc=1):
    _zstate("c", "u", Z, x, y, fold, inc)


def cbcdotc(Z, x, y, fold=SDEFAULT_FOLD, inc=1):
    _zstate("c", "c", Z, x, y, fold, inc)


def sbcasum(Y, x, fold=SDEFAULT_FOLD, inc=1):
    _zstate("c", "asum", Y, x, None, fold, inc)


def sbcssq(Y, x, scale=0.0, fold=SDEFAULT_FOLD, inc=1):
    _bind_stream(x, Y)
    r = lib().binnedBLAS_sbcssq(fold, _n_of(x, inc), _zp(x, "c"), inc, scale, _ptr(Y, "float32"))
    _check()
    return r


def _rzgemv(prec, trans, alpha, A, x, beta, y, fold, order, M, N, lda, incx, incy):
    rm = order in ("r", "R")
    if M is None or N is None or lda is None:
        sh = tuple(A.shape)
        M, N, lda = (sh[0], sh[1], sh[1]) if rm else (sh[1], sh[0], sh[1])
    al, be = _zpair(alpha, prec), _zpair(beta, prec)
    _bind_stream(A, x, y)
    getattr(lib(), f"reproBLAS_r{prec}gemv")(fold, _ch(order), _ch(trans), M, N, _ptr(al, al.dtype.name),
                                             _zp(A, prec), lda, _zp(x, prec), incx, _ptr(be, be.dtype.name),
                                             _zp(y, prec), incy)
    _check()


def _rzgemm(prec, transa, transb, alpha, A, B, beta, C, fold, order, M, N, K, lda, ldb, ldc):
    rm = order in ("r", "R")

    def dims(X):
        sh = tuple(X.shape)
        return (sh[0], sh[1], sh[1]) if rm else (sh[1], sh[0], sh[1])

    ra, ca, la = dims(A)
    rb_, cb, lb = dims(B)
    rc, cc, lc = dims(C)
    na = transa in ("n", "N")
    M = (ra if na else ca) if M is None else M
    K = (ca if na else ra) if K is None else K
    N = (cb if transb in ("n", "N") else rb_) if N is None else N
    lda, ldb, ldc = lda or la, ldb or lb, ldc or lc
    al, be = _zpair(alpha, prec), _zpair(beta, prec)
    _bind_stream(A, B, C)
    getattr(lib(), f"reproBLAS_r{prec}gemm")(fold, _ch(order), _ch(transa), _ch(transb), M, N, K,
                                             _ptr(al, al.dtype.name), _zp(A, prec), lda, _zp(B, prec), ldb,
                                             _ptr(be, be.dtype.name), _zp(C, prec), ldc)
    _check()


def rzgemv(trans, alpha, A, x, beta, y, fold=DEFAULT_FOLD, order="C", M=None, N=None, lda=None, incx=1, incy=1):
    """reproBLAS_rzgemv on a 2-D complex128 A (shapes as for :func:`rdgemv`; trans 'N', 'T' or 'C')."""
    _rzgemv("z", trans, alpha, A, x, beta, y, fold, order, M, N, lda, incx, incy)


def rzgemm(transa, transb, alpha, A, B, beta, C, fold=DEFAULT_FOLD, order="C", M=None, N=None, K=None,
           lda=None, ldb=None, ldc=None):
    """reproBLAS_rzgemm on complex128 operands; C is updated in place."""
    _rzgemm("z", transa, transb, alpha, A, B, beta, C, fold, order, M, N, K, lda, ldb, ldc)


def rcgemv(trans, alpha, A, x, beta, y, fold=SDEFAULT_FOLD, order="C", M=None, N=None, lda=None, incx=1, incy=1):
    _rzgemv("c", trans, alpha, A, x, beta, y, fold, order, M, N, lda, incx, incy)


def rcgemm(transa, transb, alpha, A, B, beta, C, fold=SDEFAULT_FOLD, order="C", M=None, N=None, K=None,
           lda=None, ldb=None, ldc=None):
    _rzgemm("c", transa, transb, alpha, A, B, beta, C, fold, order, M, N, K, lda, ldb, ldc)
