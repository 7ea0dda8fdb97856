"""Differentiable reproducible matmul: gradcheck by finite differences,
gradients bit-identical across gemm tilings, and every gemm call of forward
and backward bit-identical to the CPU reference (oracle) on the same operands
(row-major Order 'R' with the call's transpose flags)."""
import ctypes

import numpy as np
import pytest

from cpu_libs import oracle

pytestmark = pytest.mark.gpu


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_gradcheck_64(gpu):
    import torch

    from reproblas_amd.autograd import repro_matmul

    g = torch.Generator(device=gpu)
    g.manual_seed(0)
    a = torch.randn(64, 64, dtype=torch.float64, device=gpu, generator=g, requires_grad=True)
    b = torch.randn(64, 64, dtype=torch.float64, device=gpu, generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x, y: repro_matmul(x, y), (a, b), eps=1e-6, atol=1e-8, rtol=1e-6)


def _run_mlp(gpu, sizes, batch, tiling, record=None):
    import torch

    import reproblas_amd as rb
    import reproblas_amd.autograd as ag

    rb.set_gemm_tiling(tiling)
    g = torch.Generator(device=gpu)
    g.manual_seed(1)
    model = ag.ReproMLP(sizes, generator=g, device=gpu)
    x = torch.randn(batch, sizes[0], dtype=torch.float64, device=gpu, generator=g)
    t = torch.randn(batch, sizes[-1], dtype=torch.float64, device=gpu, generator=g)
    ag.RECORD = record
    try:
        loss = ag.mlp_step(model, x, t)
    finally:
        ag.RECORD = None
    return loss.detach().clone(), [w.grad.detach().clone() for w in model.weights]


def test_mlp_gradients_identical_across_tilings(gpu):
    import torch

    import reproblas_amd as rb

    old = rb.get_gemm_tiling()
    try:
        runs = [_run_mlp(gpu, (64, 128, 96, 32), 80, t) for t in (0, 1, 2)]
    finally:
        rb.set_gemm_tiling(old)
    for loss, grads in runs[1:]:
        assert torch.equal(loss.view(torch.int64), runs[0][0].view(torch.int64))
        for ga, gb in zip(grads, runs[0][1]):
            assert torch.equal(ga.view(torch.int64), gb.view(torch.int64))


def test_every_gemm_call_matches_cpu_reference(gpu):
    import reproblas_amd as rb

    rec = []
    old = rb.get_gemm_tiling()
    try:
        _run_mlp(gpu, (48, 96, 64, 24), 40, 1, record=rec)
    finally:
        rb.set_gemm_tiling(old)
    assert len(rec) == 3 + 5  # 3 forward; backward: 3 dW + 2 dX (the input needs none)
    O = oracle()
    for ta, tb, A, B, C in rec:
        A = A.cpu().numpy()
        B = B.cpu().numpy()
        Cg = C.cpu().numpy()
        m, n = Cg.shape
        k = A.shape[1] if ta == "N" else A.shape[0]
        Ce = np.zeros((m, n))
        O.fn["rdgemm"](3, b"R", ta.encode(), tb.encode(), m, n, k, 1.0, vp(np.ascontiguousarray(A)), A.shape[1],
                       vp(np.ascontiguousarray(B)), B.shape[1], 0.0, vp(Ce), n)
        assert np.array_equal(Cg.view(np.uint64), Ce.view(np.uint64)), (ta, tb, A.shape, B.shape)


@pytest.mark.slow
def test_full_mlp_identical_across_tilings(gpu):
    """Benchmark scale: 1024-4096-4096-1024, batch 2048."""
    import torch

    import reproblas_amd as rb

    old = rb.get_gemm_tiling()
    try:
        runs = [_run_mlp(gpu, (1024, 4096, 4096, 1024), 2048, t) for t in (0, 1, 2)]
    finally:
        rb.set_gemm_tiling(old)
    for loss, grads in runs[1:]:
        assert torch.equal(loss.view(torch.int64), runs[0][0].view(torch.int64))
        for ga, gb in zip(grads, runs[0][1]):
            assert torch.equal(ga.view(torch.int64), gb.view(torch.int64))
