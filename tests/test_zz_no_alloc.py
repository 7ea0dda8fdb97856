"""reproblas_hip_rddot_device_noalloc: the device dot with its scratch in the
library's fixed device array. Two calls in a row (the second smaller) and
mixed strides give the same bits as reproblas_hip_rddot_device, which the
other tests pin to the oracle; bad arguments are refused."""
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle

pytestmark = pytest.mark.gpu


def _fn(L):
    f = L.reproblas_hip_rddot_device_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                  ctypes.c_void_p, ctypes.c_void_p]
    return f


def dp(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + t.element_size() * off)


def test_noalloc_dot_matches_allocating_entry_point(gpu):
    import torch

    import reproblas_amd as rb

    L = rb.lib()
    f = _fn(L)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = (1 << 22) + 333
    x, y = dg.wide(n, 701), dg.wide(n, 702)
    x[12345] = 1.6e308  # index-0 terms in the first call ...
    y[12345] = 1.0
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    out = torch.zeros(8, dtype=torch.float64, device=gpu)
    # (fold, terms, incX, incY); the second call is smaller than the first
    cases = [(3, n, 1, 1), (3, 70001, 1, 1), (2, 4099, 2, 3), (8, 100003, 1, 1), (52, 5000, 1, 1), (30, 9000, 3, 1),
             (3, 0, 1, 1), (4, 1, 1, 1)]
    for fold, m, ix, iy in cases:
        assert (m - 1) * max(ix, iy) < n
        assert L.reproblas_hip_rddot_device(fold, m, dp(tx), ix, dp(ty), iy, dp(out, 0), s) == 0
        assert f(fold, m, dp(tx), ix, dp(ty), iy, dp(out, 1), s) == 0
        got = out[:2].cpu().numpy()
        assert got.view(np.uint64)[0] == got.view(np.uint64)[1], (fold, m, ix, iy, got)
    # the first case against the oracle directly as well
    assert f(3, n, dp(tx), 1, dp(ty), 1, dp(out, 2), s) == 0
    assert np.array([out[2].item()]).view(np.uint64)[0] == np.array([oracle().rddot(3, x, y)]).view(np.uint64)[0]


def test_noalloc_dot_infinity_replay(gpu):
    """The infinity-path replay runs inside the same launch (last workgroup),
    so it needs no scratch beyond the fixed sync block."""
    import torch

    import reproblas_amd as rb
    import test_oracle_corners as tc

    L = rb.lib()
    f = _fn(L)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.zeros(2, dtype=torch.float64, device=gpu)
    for key, op, dt, kind, n, seed, inc in tc.CASES:
        if op != "dot" or dt != np.float64 or inc != 1 or seed > 2:
            continue
        x, y = tc.make_input(op, dt, kind, n, seed, inc)
        tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
        assert f(3, n, dp(tx), 1, dp(ty), 1, dp(out), s) == 0
        assert tc.same(out[0].item(), oracle().rddot(3, x, y)), key


def test_noalloc_refuses_bad_arguments(gpu):
    import torch

    import reproblas_amd as rb

    L = rb.lib()
    f = _fn(L)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    tx = torch.zeros(16, dtype=torch.float64, device=gpu)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    assert f(1, 16, dp(tx), 1, dp(tx), 1, dp(out), s) != 0  # fold outside [2, 52]
    # strided operands use four workgroups per CU; at fold 52 that grid still
    # merges through the fixed sync block (no per-workgroup records), so the
    # call succeeds with the allocating entry point's bits
    big = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, 1 << 24)).to(gpu)
    two = torch.zeros(2, dtype=torch.float64, device=gpu)
    assert f(52, (1 << 23) - 1, dp(big), 2, dp(big, 1), 2, dp(two, 1), s) == 0
    assert L.reproblas_hip_rddot_device(52, (1 << 23) - 1, dp(big), 2, dp(big, 1), 2, dp(two, 0), s) == 0
    got = two.cpu().numpy()
    assert got.view(np.uint64)[0] == got.view(np.uint64)[1], got
    host = np.zeros(16)
    assert f(3, 16, ctypes.c_void_p(host.ctypes.data), 1, dp(tx), 1, dp(out), s) != 0  # host memory
    L.reproblas_hip_clear_error()
    assert f(3, 16, dp(tx), 1, dp(tx), 1, dp(out), s) == 0
