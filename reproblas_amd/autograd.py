"""Differentiable reproducible matmul for PyTorch (ROCm), FP64.

``repro_matmul(a, b)`` computes a @ b for row-major float64 tensors on a ROCm
device with the binned gemm (reproBLAS_rdgemm semantics, fold 3 by default):
every output element is bit-identical whatever the tiling, and to the CPU
reference library on the same operands. Its backward is binned too:
dA = dC . B^T and dB = A^T . dC, each one more binned gemm with transposed
operands, so gradients are reproducible bit for bit as well.

A torch row-major (m, k) tensor is an m x k row-major matrix, so the calls go
through Order 'R' of the C ABI (reproblas_hip_rdgemm_device) on torch's
current stream.
"""
import ctypes

import torch

from . import _check, lib

_vp = ctypes.c_void_p

# When set to a list, every gemm of the autograd path appends
# (transa, transb, A, B, C) clones here (used by the parity tests).
RECORD = None


def gemm_rowmajor(transa, transb, a, b, fold=3, out=None):
    """C = op(a) @ op(b) (row-major operands, binned fold `fold`)."""
    if a.dtype != torch.float64 or b.dtype != torch.float64:
        raise TypeError("repro matmul is FP64 only")
    if not (a.is_cuda and b.is_cuda):
        raise ValueError("repro matmul runs on the GPU; operands must be device tensors")
    a = a.contiguous()
    b = b.contiguous()
    m = a.shape[0] if transa == "N" else a.shape[1]
    k = a.shape[1] if transa == "N" else a.shape[0]
    n = b.shape[1] if transb == "N" else b.shape[0]
    kb = b.shape[0] if transb == "N" else b.shape[1]
    if k != kb:
        raise ValueError(f"inner dimensions differ: {k} vs {kb}")
    c = out if out is not None else torch.empty(m, n, dtype=torch.float64, device=a.device)
    s = torch.cuda.current_stream(a.device)
    L = lib()
    r = L.reproblas_hip_rdgemm_device(fold, b"R", transa.encode(), transb.encode(), m, n, k, 1.0, _vp(a.data_ptr()),
                                      a.shape[1], _vp(b.data_ptr()), b.shape[1], 0.0, _vp(c.data_ptr()), n,
                                      _vp(s.cuda_stream))
    if r != 0:
        _check()
    if RECORD is not None:
        RECORD.append((transa, transb, a.detach().clone(), b.detach().clone(), c.detach().clone()))
    return c


class ReproMatmul(torch.autograd.Function):
    """C = A @ B with binned forward and binned backward."""

    @staticmethod
    def forward(ctx, a, b, fold=3):
        ctx.save_for_backward(a, b)
        ctx.fold = fold
        return gemm_rowmajor("N", "N", a, b, fold)

    @staticmethod
    def backward(ctx, grad):
        a, b = ctx.saved_tensors
        grad = grad.contiguous()
        da = db = None
        if ctx.needs_input_grad[0]:
            da = gemm_rowmajor("N", "T", grad, b, ctx.fold)  # dA = dC . B^T
        if ctx.needs_input_grad[1]:
            db = gemm_rowmajor("T", "N", a, grad, ctx.fold)  # dB = A^T . dC
        return da, db, None


def repro_matmul(a, b, fold=3):
    return ReproMatmul.apply(a, b, fold)


class ReproMLP(torch.nn.Module):
    """Bias-free tanh MLP whose matmuls are all binned (FP64)."""

    def __init__(self, sizes=(1024, 4096, 4096, 1024), fold=3, generator=None, device="cuda"):
        super().__init__()
        self.fold = fold
        ws = []
        for fi, fo in zip(sizes[:-1], sizes[1:]):
            w = torch.randn(fi, fo, dtype=torch.float64, device=device, generator=generator) / (fi ** 0.5)
            ws.append(torch.nn.Parameter(w))
        self.weights = torch.nn.ParameterList(ws)

    def forward(self, x):
        h = x
        for i, w in enumerate(self.weights):
            h = repro_matmul(h, w, self.fold)
            if i + 1 < len(self.weights):
                h = torch.tanh(h)
        return h


def mlp_step(model, x, target):
    """One forward + backward (MSE loss); returns the loss tensor."""
    model.zero_grad(set_to_none=True)
    y = model(x)
    loss = torch.mean((y - target) ** 2)
    loss.backward()
    return loss
