"""Order-free exchange of binned values between processes (reproblas_amd.dist,
reproblas_hip_dbencode / dbdecode): the role of the reference's MPI merge
operator (src/binnedMPI/DBDBADD.ccog). CPU only: the wire is host memory and
the multi-process test runs torch.distributed over gloo on 127.0.0.1.

Pinned against the oracle: decoding the summed wires of any set of binned
values gives the bits of merging them one by one with binned_dbdbadd
(src/binned/dbdbadd.c, dmdmadd.c:27-74), in either order, and the
all-reduce of per-rank partial sums gives every rank the bits of one binned
sum over the whole vector, whichever rank holds which shard. Complex binned
values travel as their two parts (binned_zbzbadd merges part by part,
src/binned/zmzmadd.c)."""
import os
import socket
import tempfile

import numpy as np
import pytest

import datagen as dg
from cpu_libs import oracle, oracle_s, oracle_z


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def partials(dt, fold, seed, k=6):
    """k binned partial sums over shards of different magnitude ranges (so
    their indexes differ), one empty."""
    g = np.random.default_rng(seed)
    O = oracle() if dt == np.float64 else oracle_s()
    out = []
    for i in range(k):
        n = int(g.integers(1, 3000))
        x = (g.uniform(-1, 1, n) * np.exp2(g.integers(-30, 30) * (3 if dt == np.float64 else 1))).astype(dt)
        Y = np.zeros(2 * fold, dtype=dt)
        if i != 2:  # the third stays empty
            (O.dbdsum if dt == np.float64 else O.sbssum)(fold, x, Y)
        out.append(Y)
    return out


def merge_seq(dt, fold, vals):
    O = oracle() if dt == np.float64 else oracle_s()
    Y = np.zeros(2 * fold, dtype=dt)
    for v in vals:
        (O.dbdbadd if dt == np.float64 else O.sbsbadd)(fold, v, Y)
    return Y


@pytest.mark.parametrize("dt,fold", [(np.float64, f) for f in (2, 3, 4, 8, 20)] + [(np.float32, f) for f in (2, 3, 8)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wire_sum_equals_sequential_merges(dt, fold, seed):
    from reproblas_amd import dist

    vals = partials(dt, fold, 300 + seed)
    wires = [dist.encode(v, fold) for v in vals]
    W = np.sum([w for w, _ in wires], axis=0)
    I = min(i for _, i in wires)
    got = dist.decode(W, I, fold, dtype=dt)
    for order in (vals, vals[::-1]):
        exp = merge_seq(dt, fold, order)
        assert np.array_equal(bits(got), bits(exp)), (fold, seed)


def test_wire_classes_and_empty():
    from reproblas_amd import dist

    fold = 3
    O = oracle()
    a = np.zeros(6)
    O.dbdsum(fold, np.array([1.0, np.inf]), a)
    b = np.zeros(6)
    O.dbdsum(fold, np.array([2.0, -np.inf]), b)
    c = np.zeros(6)
    O.dbdsum(fold, np.array([3.0, 4.0]), c)
    e = np.zeros(6)
    for vals, want in (([a, c], np.inf), ([b, c], -np.inf), ([a, b, c], np.nan), ([e, e], 0.0)):
        wires = [dist.encode(v, fold) for v in vals]
        got = dist.decode(np.sum([w for w, _ in wires], axis=0), min(i for _, i in wires), fold)
        exp = merge_seq(np.float64, fold, vals)
        assert (np.isnan(got[0]) and np.isnan(exp[0])) or got[0] == exp[0] == want, (got[0], exp[0])
    assert np.array_equal(bits(dist.decode(*[np.zeros(dist.DB_WIRE, np.int64), dist.EMPTY_INDEX], fold)), bits(e))


@pytest.mark.parametrize("prec,fold", [("z", 3), ("z", 8), ("c", 3)])
def test_complex_parts_sum_equals_sequential_merges(prec, fold):
    """Summed wires of each part of complex partial sums decode to the bits
    of merging the complex values one by one with binned_zbzbadd."""
    from reproblas_amd import dist

    Z = oracle_z(prec)
    rt = np.float64 if prec == "z" else np.float32
    g = np.random.default_rng(41 + fold)
    vals = []
    for i in range(5):
        n = int(g.integers(1, 2000))
        x = (g.uniform(-1, 1, 2 * n) * np.exp2(g.integers(-20, 20))).astype(rt)
        Y = np.zeros(4 * fold, dtype=rt)
        if i != 1:
            Z.bsum(fold, x, Y)
        vals.append(Y)
    enc = [[dist.encode(p, fold) for p in dist._parts(v, fold)[1]] for v in vals]
    got = np.zeros(4 * fold, dtype=rt)
    for c in (0, 1):
        v = dist.decode(np.sum([e[c][0] for e in enc], axis=0), min(e[c][1] for e in enc), fold, dtype=rt)
        got[c:2 * fold:2] = v[:fold]
        got[2 * fold + c::2] = v[fold:]
    for order in (vals, vals[::-1]):
        exp = np.zeros(4 * fold, dtype=rt)
        for v in order:
            Z.fn["bbadd"](fold, Z.p(v), Z.p(exp))
        assert np.array_equal(bits(got), bits(exp)), (prec, fold)


def _free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def _rank_main(rank, world, port, shards, fold, out_dir):
    import torch.distributed as dist_

    from reproblas_amd import dist

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist_.init_process_group("gloo", rank=rank, world_size=world)
    try:
        Y = np.zeros(2 * fold)
        for x in shards[rank]:
            oracle().dbdsum(fold, x, Y)
        dist.allreduce_binned(Y, fold)
        np.save(os.path.join(out_dir, f"r{rank}.npy"), Y)
    finally:
        dist_.destroy_process_group()


@pytest.mark.parametrize("world", [2, 3])
def test_allreduce_over_gloo_matches_one_sum(world):
    import torch.multiprocessing as mp

    fold = 3
    x = dg.wide(40001, 77)
    x[123] = 1.0e300  # one shard raises the index of the whole sum
    pieces = np.array_split(x, 6)
    O = oracle()
    full = np.zeros(2 * fold)
    O.dbdsum(fold, x, full)
    for perm in (list(range(6)), [5, 2, 4, 0, 3, 1]):
        shards = [[np.ascontiguousarray(pieces[p]) for p in perm[r::world]] for r in range(world)]
        with tempfile.TemporaryDirectory() as d:
            mp.spawn(_rank_main, args=(world, _free_port(), shards, fold, d), nprocs=world, join=True)
            got = [np.load(os.path.join(d, f"r{r}.npy")) for r in range(world)]
        for r in range(world):
            assert np.array_equal(bits(got[r]), bits(full)), (world, perm, r)
        assert O.ddbconv(fold, got[0]) == O.rdsum(fold, x)


def _rank_main_z(rank, world, port, shards, fold, out_dir):
    import torch.distributed as dist_

    from reproblas_amd import dist

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist_.init_process_group("gloo", rank=rank, world_size=world)
    try:
        Y = np.zeros(4 * fold)
        for x in shards[rank]:
            oracle_z("z").bsum(fold, x, Y)
        dist.allreduce_binned_complex(Y, fold)
        np.save(os.path.join(out_dir, f"r{rank}.npy"), Y)
    finally:
        dist_.destroy_process_group()


def test_complex_allreduce_over_gloo_matches_one_sum():
    import torch.multiprocessing as mp

    fold, world = 3, 2
    g = np.random.default_rng(5)
    x = (g.uniform(-1, 1, 2 * 9000) * np.exp2(g.integers(-40, 40, 2 * 9000))).astype(np.float64)
    x[2 * 777 + 1] = 3.0e250  # the imaginary part of one shard raises that part's index
    pieces = [np.ascontiguousarray(p) for p in np.split(x, 6)]  # 1500 complex each
    Z = oracle_z("z")
    full = np.zeros(4 * fold)
    Z.bsum(fold, x, full)
    shards = [pieces[r::world] for r in range(world)]
    with tempfile.TemporaryDirectory() as d:
        mp.spawn(_rank_main_z, args=(world, _free_port(), shards, fold, d), nprocs=world, join=True)
        got = [np.load(os.path.join(d, f"r{r}.npy")) for r in range(world)]
    for r in range(world):
        assert np.array_equal(bits(got[r]), bits(full)), r


def ssq_partials(dt, fold, seed, k=5):
    """k scaled sums of squares over shards of different magnitude ranges
    (normal range: every rescale exact), one empty."""
    g = np.random.default_rng(seed)
    O = oracle() if dt == np.float64 else oracle_s()
    out = []
    for i in range(k):
        n = int(g.integers(1, 3000))
        x = (g.uniform(-1, 1, n) * np.exp2(g.integers(-30, 30) * (2 if dt == np.float64 else 1))).astype(dt)
        Y = np.zeros(2 * fold, dtype=dt)
        s = 0.0
        if i != 3:
            s = (O.dbdssq if dt == np.float64 else O.sbsssq)(fold, x, 0.0, Y)
        out.append((Y, dt(s)))
    return out


@pytest.mark.parametrize("dt,fold", [(np.float64, 3), (np.float64, 8), (np.float32, 3)])
def test_ssq_exchange_equals_sequential_scaled_merges(dt, fold):
    """Rescaled to the largest scale and summed as wires: the bits and scale
    of merging with binned_dbdbaddsq / sbsbaddsq one by one, in either order."""
    import ctypes as ct

    from reproblas_amd import dist

    vp = lambda a: ct.c_void_p(a.ctypes.data)
    O = oracle() if dt == np.float64 else oracle_s()
    addsq = O.fn["dbdbaddsq" if dt == np.float64 else "sbsbaddsq"]
    vals = ssq_partials(dt, fold, 900 + fold)
    S = max(s for _, s in vals)
    wires = [dist.encode(dist.rescale(Y.copy(), s, S, fold), fold) for Y, s in vals]
    got = dist.decode(np.sum([w for w, _ in wires], axis=0), min(i for _, i in wires), fold, dtype=dt)
    for order in (vals, vals[::-1]):
        E, se = np.zeros(2 * fold, dtype=dt), dt(0.0)
        for Y, s in order:
            se = dt(addsq(fold, s, vp(Y), se, vp(E)))
        assert se == S
        assert np.array_equal(bits(got), bits(E)), (fold, dt)


def _rank_main_ssq(rank, world, port, shards, fold, out_dir):
    import torch.distributed as dist_

    from reproblas_amd import dist

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist_.init_process_group("gloo", rank=rank, world_size=world)
    try:
        Y, s = np.zeros(2 * fold), 0.0
        for x in shards[rank]:
            s = oracle().dbdssq(fold, x, s, Y)
        s = dist.allreduce_binned_ssq(Y, s, fold)
        np.save(os.path.join(out_dir, f"r{rank}.npy"), np.concatenate([Y, [s]]))
    finally:
        dist_.destroy_process_group()


def test_ssq_allreduce_over_gloo_matches_sequential_merges():
    """Each rank's dbdssq over its shards, exchanged: every rank gets the
    bits and scale of merging the ranks' values with dbdbaddsq."""
    import ctypes as ct
    import torch.multiprocessing as mp

    fold, world = 3, 2
    g = np.random.default_rng(17)
    pieces = [(g.uniform(-1, 1, 4000) * np.exp2(e)).astype(np.float64) for e in (-60, 5, 80, -10)]
    shards = [pieces[r::world] for r in range(world)]
    O = oracle()
    vp = lambda a: ct.c_void_p(a.ctypes.data)
    E, se = np.zeros(2 * fold), 0.0
    for r in range(world):
        Y, s = np.zeros(2 * fold), 0.0
        for x in shards[r]:
            s = O.dbdssq(fold, x, s, Y)
        se = O.fn["dbdbaddsq"](fold, s, vp(Y), se, vp(E))
    with tempfile.TemporaryDirectory() as d:
        mp.spawn(_rank_main_ssq, args=(world, _free_port(), shards, fold, d), nprocs=world, join=True)
        got = [np.load(os.path.join(d, f"r{r}.npy")) for r in range(world)]
    for r in range(world):
        assert got[r][-1] == se, r
        assert np.array_equal(bits(got[r][:-1]), bits(E)), r
