"""Ragged batches of reproducible reductions and gemms in one launch.

    b = Batch(fold=3)
    b.add_segments("dot", offsets, x, y, results)        # CSR segments (torch, device)
    b.add_gemms("N", "N", [(A0, B0, C0), ...], alpha=1.0, beta=0.0)
    b.launch(graph=True)                                  # one hipGraph replay

Each segment result is bit-identical to reproBLAS_rddot (or rdsum/rdasum/
rdnrm2) on that segment, each gemm to reproBLAS_rdgemm (column-major,
like the reference's default Order). Plans are built once (host-side tables,
device allocations) and can be launched any number of times.
"""
import ctypes

from . import _ch, _check, _ptr, lib

_OPS = {"sum": 0, "asum": 1, "dot": 2, "nrm2": 3}
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


def _sigs(L):
    if getattr(L, "_batch_sigs", False):
        return
    L.reproblas_hip_batch_create.restype = _vp
    L.reproblas_hip_batch_create.argtypes = [ctypes.c_int]
    L.reproblas_hip_batch_add_segments.restype = ctypes.c_int
    L.reproblas_hip_batch_add_segments.argtypes = [_vp, ctypes.c_int, _i64, ctypes.POINTER(_i64), _vp, _vp, _vp,
                                                   _i64]
    P64 = ctypes.POINTER(_i64)
    PD = ctypes.POINTER(ctypes.c_double)
    PP = ctypes.POINTER(_vp)
    L.reproblas_hip_batch_add_gemms.restype = ctypes.c_int
    L.reproblas_hip_batch_add_gemms.argtypes = [_vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int, P64,
                                                P64, P64, PD, PP, P64, PP, P64, PD, PP, P64, ctypes.c_int]
    L.reproblas_hip_batch_launch.restype = ctypes.c_int
    L.reproblas_hip_batch_launch.argtypes = [_vp, _vp]
    L.reproblas_hip_batch_graph_launch.restype = ctypes.c_int
    L.reproblas_hip_batch_graph_launch.argtypes = [_vp, _vp]
    L.reproblas_hip_batch_destroy.restype = None
    L.reproblas_hip_batch_destroy.argtypes = [_vp]
    L._batch_sigs = True


class Batch:
    """A plan of ragged segment reductions and small gemms (device operands)."""

    def __init__(self, fold=3):
        self.L = lib()
        _sigs(self.L)
        self.fold = fold
        self.h = self.L.reproblas_hip_batch_create(fold)
        if not self.h:
            _check()
        self._keep = []  # operands referenced by the plan stay alive with it

    def add_segments(self, op, offsets, x, y=None, results=None, chunk=0):
        """offsets: host sequence of nseg+1 int64 (CSR); x, y, results: device tensors."""
        import numpy as np

        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        nseg = len(off) - 1
        r = self.L.reproblas_hip_batch_add_segments(self.h, _OPS[op], nseg, off.ctypes.data_as(ctypes.POINTER(_i64)),
                                                    _ptr(x), _ptr(y), _ptr(results), chunk)
        if r != 0:
            _check()
        self._keep += [x, y, results]

    def add_gemms(self, transa, transb, triples, alpha=1.0, beta=0.0, tiling=-1):
        """triples: list of (A, B, C) column-major device tensors viewed as in
        reproblas_amd.rdgemm (a C-contiguous (cols, rows) tensor is a rows x
        cols column-major matrix)."""
        n = len(triples)
        M = (_i64 * n)()
        N = (_i64 * n)()
        K = (_i64 * n)()
        lda = (_i64 * n)()
        ldb = (_i64 * n)()
        ldc = (_i64 * n)()
        al = (ctypes.c_double * n)()
        be = (ctypes.c_double * n)()
        Ap = (_vp * n)()
        Bp = (_vp * n)()
        Cp = (_vp * n)()
        na = transa in ("n", "N")
        nb = transb in ("n", "N")
        for i, (A, B, C) in enumerate(triples):
            ar, ac = A.shape[1], A.shape[0]
            br, bc = B.shape[1], B.shape[0]
            M[i] = ar if na else ac
            K[i] = ac if na else ar
            N[i] = bc if nb else br
            lda[i], ldb[i], ldc[i] = ar, br, C.shape[1]
            al[i] = alpha[i] if isinstance(alpha, (list, tuple)) else alpha
            be[i] = beta[i] if isinstance(beta, (list, tuple)) else beta
            Ap[i], Bp[i], Cp[i] = _ptr(A).value, _ptr(B).value, _ptr(C).value
        r = self.L.reproblas_hip_batch_add_gemms(self.h, b"C", _ch(transa), _ch(transb), n, M, N, K, al, Ap, lda, Bp,
                                                 ldb, be, Cp, ldc, tiling)
        if r != 0:
            _check()
        self._keep += list(triples)

    def launch(self, graph=True, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        fn = self.L.reproblas_hip_batch_graph_launch if graph else self.L.reproblas_hip_batch_launch
        if fn(self.h, _vp(s.cuda_stream)) != 0:
            _check()

    def close(self):
        if getattr(self, "h", None):
            self.L.reproblas_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
