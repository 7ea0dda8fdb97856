"""GPU parity in the regions where the reference's result depends on its
block schedule (the oracle is pinned there against the compiled reference
and tests/golden/corners.npz by tests/test_oracle_corners.py):

* the infinity path of rdsum / rddot / rssum / rsdot (dmdsum.ccog:66-78,
  dmddot.ccog:68-81): a sequential overflow before the first infinity of the
  reference's first infinite block makes the result NaN;
* nrm2 over values spanning 2^-1000..2^1000 with the largest values in late
  blocks (dmdssq.ccog:60-101), at every fold.

Bit for bit against the oracle (NaN compared by class), host and device
operands, unit and non-unit strides.
"""
import ctypes

import numpy as np
import pytest

import test_oracle_corners as tc
from cpu_libs import oracle, oracle_s, oracle_z

pytestmark = pytest.mark.gpu


def gpu_run(rb, op, dt, fold, x, y, inc, device):
    import torch

    if device is not None:
        x = torch.from_numpy(x).to(device)
        y = torch.from_numpy(y).to(device) if y is not None else None
    if dt == np.float64:
        return rb.rdsum(x, fold, inc) if op == "sum" else rb.rddot(x, y, fold, inc, inc)
    return rb.rssum(x, fold, inc) if op == "sum" else rb.rsdot(x, y, fold, inc, inc)


@pytest.mark.parametrize("key,op,dt,kind,n,seed,inc", tc.CASES, ids=[c[0] for c in tc.CASES])
def test_infinity_path(gpu, key, op, dt, kind, n, seed, inc):
    import reproblas_amd as rb

    x, y = tc.make_input(op, dt, kind, n, seed, inc)
    orc = oracle() if dt == np.float64 else oracle_s()
    eq = tc.same if dt == np.float64 else tc.same32
    for fold in (2, 3, 4):
        e = tc.run(orc, op, dt, fold, x, y, inc)
        for device in (None, gpu):
            r = gpu_run(rb, op, dt, fold, x, y, inc, device)
            assert eq(r, e), (key, fold, "device" if device is not None else "host", r, e)


def test_infinity_path_accumulate_into_top_state(gpu):
    """binnedBLAS_dbdsum into a caller value at index 0, host and device Y."""
    import torch

    import reproblas_amd as rb
    import datagen as dg

    O = oracle()
    for fold in (2, 3, 4):
        Y0 = np.zeros(2 * fold)
        O.dbdsum(fold, np.array([1.6e308, 1.0]), Y0)
        x = dg.uniform(5000, 9)
        x[3000], x[3001] = 1.0e308, -np.inf
        Ye = Y0.copy()
        O.dbdsum(fold, x, Ye)
        Yh = Y0.copy()
        rb.dbdsum(Yh, x, fold)
        assert tc.same(Yh[0], Ye[0]), fold
        Yd = torch.from_numpy(Y0.copy()).to(gpu)
        rb.dbdsum(Yd, torch.from_numpy(x).to(gpu), fold)
        assert tc.same(Yd[0].item(), Ye[0]), fold


@pytest.mark.parametrize("seed,fold", tc.nrm2_range_cases())
def test_nrm2_wide_range(gpu, seed, fold):
    """rdnrm2 and the dbdssq state over a 2^2000 dynamic range whose largest
    values arrive in late blocks (the reference rescales block by block)."""
    import torch

    import reproblas_amd as rb

    x = tc.nrm2_range_input(seed)
    O = oracle()
    e = O.rdnrm2(fold, x)
    tx = torch.from_numpy(x).to(gpu)
    assert tc.same(rb.rdnrm2(tx, fold), e), (seed, fold)
    assert tc.same(rb.rdnrm2(x, fold), e), (seed, fold)
    Ye = np.zeros(2 * fold)
    se = O.dbdssq(fold, x, 0.0, Ye)
    Yg = np.zeros(2 * fold)
    sg = rb.dbdssq(Yg, x, 0.0, fold)
    assert tc.same(sg, se)
    # every slot inside the bin table; past it (folds >= 28 here) the
    # reference's own content depends on how its calls split the terms
    # (test_oracle_corners.test_ssq_slots_beyond_bin_table_depend_on_the_split)
    keep = tc.table_slots(fold, Ye, O)
    if fold <= 27:
        assert len(keep) == 2 * fold
    assert np.array_equal(Yg[keep].view(np.uint64), Ye[keep].view(np.uint64)), (seed, fold)
    assert tc.same(O.ddbconv(fold, Yg), O.ddbconv(fold, Ye))


@pytest.mark.parametrize("name,x,y", tc.AMAX_CASES, ids=[c[0] for c in tc.AMAX_CASES])
def test_amax_nan_semantics(gpu, name, x, y):
    """binnedBLAS_damax / damaxm / samax / samaxm with NaN terms: the
    reference's scalar answer (max after the last NaN term, or NaN when it is
    last), host and device operands, against the oracle."""
    import torch

    import reproblas_amd as rb

    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    n = len(x)
    O, Os = oracle(), oracle_s()
    e = O.fn["damaxm"](n, vp(x), 1, vp(y), 1) if y is not None else O.fn["damax"](n, vp(x), 1)
    xs = x.astype(np.float32)
    ys = y.astype(np.float32) if y is not None else None
    es = Os.fn["samaxm"](n, vp(xs), 1, vp(ys), 1) if ys is not None else Os.fn["samax"](n, vp(xs), 1)
    for dev in (None, gpu):
        tx = x if dev is None else torch.from_numpy(x).to(dev)
        txs = xs if dev is None else torch.from_numpy(xs).to(dev)
        if y is None:
            g, gs = rb.damax(tx), rb.samax(txs)
        else:
            ty = y if dev is None else torch.from_numpy(y).to(dev)
            tys = ys if dev is None else torch.from_numpy(ys).to(dev)
            g, gs = rb.damaxm(tx, ty), rb.samaxm(txs, tys)
        assert tc.same(g, e), (name, dev, g, e)
        assert tc.same32(gs, es), (name, dev, gs, es)


@pytest.mark.parametrize("dt,seed,fold", tc.SSQ_SUB_CASES,
                         ids=[f"{'d' if c[0] == np.float64 else 's'}{c[1]}_f{c[2]}" for c in tc.SSQ_SUB_CASES])
def test_ssq_subnormal_squares(gpu, dt, seed, fold):
    """dbdssq / sbsssq state (every slot), scale and nrm2 where scaled squares
    are subnormal at deep folds: the reference's one-call block schedule
    (dmdssq.ccog:60-101), host and device operands."""
    import torch

    import reproblas_amd as rb

    x = tc.ssq_subnormal_input(dt, seed)
    d = dt == np.float64
    O = oracle() if d else oracle_s()
    eq = tc.same if d else tc.same32
    u = np.uint64 if d else np.uint32
    Ye = np.zeros(2 * fold, dtype=dt)
    se = O.dbdssq(fold, x, 0.0, Ye) if d else O.sbsssq(fold, x, 0.0, Ye)
    Ye2 = Ye.copy()
    se2 = O.dbdssq(fold, x, se, Ye2) if d else O.sbsssq(fold, x, se, Ye2)
    ne = O.rdnrm2(fold, x) if d else O.rsnrm2(fold, x)
    for dev in (None, gpu):
        tx = x if dev is None else torch.from_numpy(x).to(dev)
        Yg = np.zeros(2 * fold, dtype=dt)
        sg = rb.dbdssq(Yg, tx, 0.0, fold) if d else rb.sbsssq(Yg, tx, 0.0, fold)
        assert eq(sg, se), (seed, fold, dev)
        assert np.array_equal(Yg.view(u), Ye.view(u)), (seed, fold, dev, np.flatnonzero(Yg.view(u) != Ye.view(u)))
        # the same terms accumulated into the result (a caller's value)
        sg2 = rb.dbdssq(Yg, tx, sg, fold) if d else rb.sbsssq(Yg, tx, sg, fold)
        assert eq(sg2, se2), (seed, fold, dev)
        assert np.array_equal(Yg.view(u), Ye2.view(u)), (seed, fold, dev, np.flatnonzero(Yg.view(u) != Ye2.view(u)))
        assert eq(rb.rdnrm2(tx, fold) if d else rb.rsnrm2(tx, fold), ne), (seed, fold, dev)


@pytest.mark.parametrize("prec,seed,blocks,fold", tc.ZSSQ_SUB_CASES, ids=tc.ZSSQ_SUB_IDS)
def test_zssq_subnormal_squares(gpu, prec, seed, blocks, fold):
    """dbzssq / sbcssq state (every slot) and scale, and dznrm2 / scnrm2,
    where scaled squares are subnormal at deep folds: the reference's
    one-call block schedule (dmzssq.ccog:59-135), host and device operands,
    then the same terms accumulated into the result (a caller's value)."""
    import torch

    import reproblas_amd as rb

    x = tc.zssq_subnormal_input(prec, seed, blocks)
    z = prec == "z"
    cx = x.view(np.complex128 if z else np.complex64)
    Z = oracle_z(prec)
    u = np.uint64 if z else np.uint32
    Ye = np.zeros(2 * fold, dtype=x.dtype)
    se = Z.bssq(fold, x, 0.0, Ye)
    Ye2 = Ye.copy()
    se2 = Z.bssq(fold, x, se, Ye2)
    ne = np.array([Z.rnrm2(fold, x)], x.dtype)
    ssq = rb.dbzssq if z else rb.sbcssq
    nrm2 = rb.rdznrm2 if z else rb.rscnrm2
    for dev in (None, gpu):
        tx = cx if dev is None else torch.from_numpy(cx).to(dev)
        Yg = np.zeros(2 * fold, dtype=x.dtype)
        sg = ssq(Yg, tx, 0.0, fold)
        assert np.array_equal(np.array([sg], x.dtype).view(u), np.array([se], x.dtype).view(u)), (dev, sg, se)
        assert np.array_equal(Yg.view(u), Ye.view(u)), (dev, np.flatnonzero(Yg.view(u) != Ye.view(u)))
        sg2 = ssq(Yg, tx, sg, fold)
        assert np.array_equal(np.array([sg2], x.dtype).view(u), np.array([se2], x.dtype).view(u)), (dev, sg2, se2)
        assert np.array_equal(Yg.view(u), Ye2.view(u)), (dev, np.flatnonzero(Yg.view(u) != Ye2.view(u)))
        assert np.array_equal(np.array([nrm2(tx, fold)], x.dtype).view(u), ne.view(u)), dev


@pytest.mark.parametrize("seed,fold", [(seed, fold) for seed in range(2) for fold in (3, 12, 20, 27, 40, 52)])
def test_znrm2_wide_range(gpu, seed, fold):
    """dznrm2 and the dbzssq state over the same 2^2000 range read as complex
    (the largest values in late 1024-element blocks, so the reference's
    scale rises block by block): scale, nrm2 and every slot inside the bin
    table; past it (folds >= 28 here) the reference's own content depends on
    its schedule, as for the real sum of squares."""
    import torch

    import reproblas_amd as rb

    x = tc.nrm2_range_input(seed)
    x = np.ascontiguousarray(x[: len(x) // 2 * 2])
    cx = x.view(np.complex128)
    Z, O = oracle_z("z"), oracle()
    Ye = np.zeros(2 * fold)
    se = Z.bssq(fold, x, 0.0, Ye)
    ne = Z.rnrm2(fold, x)
    keep = tc.table_slots(fold, Ye, O)
    if fold <= 27:
        assert len(keep) == 2 * fold
    for dev in (None, gpu):
        tx = cx if dev is None else torch.from_numpy(cx).to(dev)
        Yg = np.zeros(2 * fold)
        sg = rb.dbzssq(Yg, tx, 0.0, fold)
        assert tc.same(sg, se), (seed, fold, dev)
        assert np.array_equal(Yg[keep].view(np.uint64), Ye[keep].view(np.uint64)), (seed, fold, dev)
        assert tc.same(O.ddbconv(fold, Yg), O.ddbconv(fold, Ye)), (seed, fold, dev)
        assert tc.same(rb.rdznrm2(tx, fold), ne), (seed, fold, dev)
