"""Ragged batches (one persistent launch per group, hipGraph replay) and the
autotuner: every result equals the single call's bits (oracle), for every
tuning knob, on every graph replay."""
import ctypes

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle

pytestmark = pytest.mark.gpu


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def bits(t):
    return t.detach().cpu().numpy().view(np.uint64)


def ragged_offsets(nseg, seed, lo=16, hi=1 << 14, with_empty=True):
    g = dg.rng(seed)
    lens = np.exp(g.uniform(np.log(lo), np.log(hi), nseg)).astype(np.int64)
    if with_empty:
        lens[g.choice(nseg, max(1, nseg // 20), replace=False)] = 0
    lens[0] = 1
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


@pytest.mark.parametrize("op", ["sum", "asum", "dot", "nrm2"])
@pytest.mark.parametrize("chunk", [0, 256, 4096])
def test_segments_match_single_calls(gpu, op, chunk):
    import torch

    from reproblas_amd.batch import Batch

    off = ragged_offsets(300, 7)
    n = int(off[-1])
    x = dg.wide(n, 1)
    y = dg.wide(n, 2)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    res = torch.full((len(off) - 1,), -1.0, dtype=torch.float64, device=gpu)
    b = Batch(3)
    b.add_segments(op, off, tx, ty if op == "dot" else None, res, chunk=chunk)
    O = oracle()
    exp = []
    for s in range(len(off) - 1):
        xs, ys = x[off[s]:off[s + 1]], y[off[s]:off[s + 1]]
        exp.append({"sum": O.rdsum, "asum": O.rdasum, "nrm2": O.rdnrm2}[op](3, xs) if op != "dot"
                   else O.rddot(3, xs, ys))
    exp = np.array(exp)
    b.launch(graph=False)
    assert np.array_equal(bits(res), exp.view(np.uint64))
    for rep in range(3):  # graph replays, with eager work in between
        res.fill_(-1.0)
        _ = torch.sum(tx * 2.0)
        torch.cuda.synchronize()
        b.launch(graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(bits(res), exp.view(np.uint64)), rep
    b.close()


def test_segments_special_values(gpu):
    import torch

    from reproblas_amd.batch import Batch

    off = np.array([0, 5, 900, 901, 5000, 5000, 9000], dtype=np.int64)
    x = dg.uniform(9000, 3)
    x[2] = np.nan
    x[100] = np.inf
    x[300] = -np.inf
    x[900] = np.inf
    x[6000] = -np.inf
    tx = torch.from_numpy(x).to(gpu)
    res = torch.zeros(6, dtype=torch.float64, device=gpu)
    b = Batch(3)
    b.add_segments("sum", off, tx, None, res)
    b.launch(graph=False)
    torch.cuda.synchronize()
    got = res.cpu().numpy()
    O = oracle()
    for s in range(6):
        e = O.rdsum(3, np.ascontiguousarray(x[off[s]:off[s + 1]]))
        assert (np.isnan(e) and np.isnan(got[s])) or got[s] == e, s


def _gemm_problem(count, seed, dev, beta_nonzero=False):
    import torch

    g = dg.rng(seed)
    sizes = g.integers(1, 9, (count, 3)) * 32
    triples, host = [], []
    for M, N, K in sizes:
        A = dg.normal((K, M), int(g.integers(1 << 30)))
        B = dg.normal((N, K), int(g.integers(1 << 30)))
        C = dg.normal((N, M), int(g.integers(1 << 30))) if beta_nonzero else np.zeros((N, M))
        host.append((A, B, C))
        triples.append((torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), torch.from_numpy(C).to(dev)))
    return sizes, triples, host


@pytest.mark.parametrize("tiling", [0, 1, 2])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_gemm_batch_matches_single_calls(gpu, tiling, beta):
    import torch

    from reproblas_amd.batch import Batch

    sizes, triples, host = _gemm_problem(12, 11, gpu, beta != 0.0)
    O = oracle()
    exp = []
    for (M, N, K), (A, B, C) in zip(sizes, host):
        Ce = C.copy()
        O.fn["rdgemm"](3, b"C", b"N", b"N", int(M), int(N), int(K), 1.5, vp(A), int(M), vp(B), int(K), beta, vp(Ce),
                       int(M))
        exp.append(Ce)
    b = Batch(3)
    b.add_gemms("N", "N", triples, alpha=1.5, beta=beta, tiling=tiling)
    for graph in (False, True, True):
        for (_, _, tC), (_, _, C) in zip(triples, host):
            tC.copy_(torch.from_numpy(C))
        torch.cuda.synchronize()
        b.launch(graph=graph)
        torch.cuda.synchronize()
        for (_, _, tC), Ce in zip(triples, exp):
            assert np.array_equal(bits(tC), Ce.view(np.uint64))
    b.close()


@pytest.mark.parametrize("tiling", [1, 2])
def test_mixed_plan_segments_and_gemms(gpu, tiling):
    """A plan with segment groups and a gemm group runs the gemms on a forked
    branch (a side stream, a parallel graph branch): every result still equals
    its single call, on direct launches and on graph replays."""
    import torch

    from reproblas_amd.batch import Batch

    off = ragged_offsets(200, 21)
    n = int(off[-1])
    x = dg.wide(n, 5)
    y = dg.wide(n, 6)
    tx, ty = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    rdot = torch.full((len(off) - 1,), -1.0, dtype=torch.float64, device=gpu)
    rsum = torch.full((len(off) - 1,), -1.0, dtype=torch.float64, device=gpu)
    sizes, triples, host = _gemm_problem(10, 23, gpu, True)
    O = oracle()
    exp_dot = np.array([O.rddot(3, x[off[s]:off[s + 1]], y[off[s]:off[s + 1]]) for s in range(len(off) - 1)])
    exp_sum = np.array([O.rdsum(3, np.ascontiguousarray(x[off[s]:off[s + 1]])) for s in range(len(off) - 1)])
    exp_c = []
    for (M, N, K), (A, B, C) in zip(sizes, host):
        Ce = C.copy()
        O.fn["rdgemm"](3, b"C", b"N", b"N", int(M), int(N), int(K), 0.75, vp(A), int(M), vp(B), int(K), 2.0,
                       vp(Ce), int(M))
        exp_c.append(Ce)
    b = Batch(3)
    b.add_segments("dot", off, tx, ty, rdot)
    b.add_gemms("N", "N", triples, alpha=0.75, beta=2.0, tiling=tiling)
    b.add_segments("sum", off, tx, None, rsum, chunk=1024)
    for rep, graph in enumerate((False, True, True, True, False)):
        rdot.fill_(-1.0)
        rsum.fill_(-1.0)
        for (_, _, tC), (_, _, C) in zip(triples, host):
            tC.copy_(torch.from_numpy(C))
        _ = torch.sum(tx * 2.0)  # eager work between replays
        torch.cuda.synchronize()
        b.launch(graph=graph)
        torch.cuda.synchronize()
        assert np.array_equal(bits(rdot), exp_dot.view(np.uint64)), rep
        assert np.array_equal(bits(rsum), exp_sum.view(np.uint64)), rep
        for (_, _, tC), Ce in zip(triples, exp_c):
            assert np.array_equal(bits(tC), Ce.view(np.uint64)), rep
    b.close()


def test_autotuner_picks_a_valid_config(gpu, tmp_path, monkeypatch):
    import torch

    import reproblas_amd.autotune as at

    monkeypatch.setattr(at, "CACHE", str(tmp_path / "tune.json"))
    _, triples, _ = _gemm_problem(6, 21, gpu)
    rec = at.tune_gemm_batch("N", "N", triples)
    assert rec["tiling"] in (0, 1, 2)
    off = ragged_offsets(50, 5)
    x = torch.from_numpy(dg.wide(int(off[-1]), 9)).to(gpu)
    rec2 = at.tune_segments("sum", off, x, candidates=(512, 4096))
    assert rec2["chunk"] in (512, 4096)
    # cached on second call
    assert at.tune_gemm_batch("N", "N", triples) == rec
