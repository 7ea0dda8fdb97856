"""Reproducible reduction of binned values across processes (torch.distributed).

The reference reduces binned values across MPI ranks with a user-defined MPI
operator that merges pairs in whatever order MPI picks
(src/binnedMPI/DBDBADD.ccog, binnedMPI_DBDBADD); its bits hold because the
merge is exact. Here every rank encodes its value as a "wire" of int64 words
(reproblas_hip_dbencode: the offsets of its primaries from their bins and its
carries at their absolute bin indexes, plus class counts) and its index; one
all-reduce SUM of the wires and one all-reduce MIN of the indexes, decoded on
every rank (reproblas_hip_dbdecode), give every rank the bits binned_dbdbadd
gives merging the values in any order. Integer addition is associative, so
the ring or tree order of the collective does not matter.

With the gloo backend the tensors are host tensors; with nccl (RCCL) they
would be device tensors on the rank's GPU. The wire is 211 words for double
(1.7 KB), 87 for float, whatever the fold.

A complex binned value (double_complex_binned / float_complex_binned: the
primaries of the real and imaginary parts interleaved, then the carries
likewise) is two real binned values, merged component by component by the
reference (binned_zbzbadd = dmdmadd on each part, src/binned/zmzmadd.c); it
travels as two wires and two indexes in the same two all-reduces.
"""
import ctypes

import numpy as np

from . import lib

DB_WIRE = 2 * 104 + 3
SB_WIRE = 2 * 42 + 3
EMPTY_INDEX = 1 << 20


def _fns():
    L = lib()
    out = {}
    for name, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
        enc = getattr(L, f"reproblas_hip_{name}bencode")
        enc.restype = ctypes.c_int
        enc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        dec = getattr(L, f"reproblas_hip_{name}bdecode")
        dec.restype = ctypes.c_int
        dec.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        out[name] = (enc, dec)
    return out


def encode(Y, fold):
    """(wire, index) of the binned value Y (numpy float64 or float32 array of
    2*fold elements)."""
    Y = np.ascontiguousarray(Y)
    kind = "d" if Y.dtype == np.float64 else "s"
    if Y.dtype not in (np.float64, np.float32) or Y.size != 2 * fold:
        raise ValueError("Y must hold 2*fold float64 or float32 elements")
    wire = np.zeros(DB_WIRE if kind == "d" else SB_WIRE, dtype=np.int64)
    index = np.zeros(1, dtype=np.int64)
    enc, _ = _fns()[kind]
    rc = enc(fold, Y.ctypes.data, wire.ctypes.data, index.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"reproblas_hip_{kind}bencode failed ({rc}): {lib().reproblas_hip_last_error_string()}")
    return wire, int(index[0])


def decode(wire, index, fold, dtype=np.float64):
    """The binned value (2*fold elements of dtype) of a summed wire and the
    minimum index."""
    kind = "d" if dtype == np.float64 else "s"
    wire = np.ascontiguousarray(wire, dtype=np.int64)
    if wire.size != (DB_WIRE if kind == "d" else SB_WIRE):
        raise ValueError("wire has the wrong length")
    Y = np.zeros(2 * fold, dtype=dtype)
    _, dec = _fns()[kind]
    rc = dec(fold, wire.ctypes.data, int(index), Y.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"reproblas_hip_{kind}bdecode failed ({rc}): {lib().reproblas_hip_last_error_string()}")
    return Y


def allreduce_binned(Y, fold, group=None):
    """Replaces this rank's binned value Y (numpy array, 2*fold float64 or
    float32) by the merge of every rank's value, bit-identical on every rank
    and for any rank order. Collective: every rank of `group` calls it."""
    import torch
    import torch.distributed as dist

    wire, index = encode(Y, fold)
    tw = torch.from_numpy(wire)
    ti = torch.tensor([index], dtype=torch.int64)
    dist.all_reduce(tw, op=dist.ReduceOp.SUM, group=group)
    dist.all_reduce(ti, op=dist.ReduceOp.MIN, group=group)
    Y[...] = decode(tw.numpy(), int(ti.item()), fold, dtype=Y.dtype)
    return Y


def _parts(Y, fold):
    """The real and imaginary binned parts (views into Y as a real array of
    4*fold elements) of a complex binned value."""
    R = Y.view(np.float64 if Y.dtype in (np.complex128, np.float64) else np.float32).reshape(-1)
    if R.size != 4 * fold:
        raise ValueError("a complex binned value holds 4*fold reals (2*fold complex)")
    return R, [np.concatenate([R[c:2 * fold:2], R[2 * fold + c::2]]) for c in (0, 1)]


def allreduce_binned_complex(Y, fold, group=None):
    """Replaces this rank's complex binned value Y (numpy complex128 / complex64
    array of 2*fold elements, or the real view of 4*fold) by the merge of every
    rank's value, bit-identical on every rank and for any rank order: each
    part as allreduce_binned merges a real value. Collective."""
    import torch
    import torch.distributed as dist

    R, parts = _parts(Y, fold)
    enc = [encode(p, fold) for p in parts]
    tw = torch.from_numpy(np.concatenate([w for w, _ in enc]))
    ti = torch.tensor([i for _, i in enc], dtype=torch.int64)
    dist.all_reduce(tw, op=dist.ReduceOp.SUM, group=group)
    dist.all_reduce(ti, op=dist.ReduceOp.MIN, group=group)
    W = tw.numpy().reshape(2, -1)
    for c in (0, 1):
        v = decode(W[c], int(ti[c].item()), fold, dtype=R.dtype)
        R[c:2 * fold:2] = v[:fold]
        R[2 * fold + c::2] = v[fold:]
    return Y


def rescale(Y, scale, new_scale, fold):
    """Y (a scaled binned sum of squares at `scale`, a contiguous numpy array
    of 2*fold float64 or float32) re-expressed at new_scale >= scale:
    binned_dmdrescale / binned_smsrescale in place."""
    if Y.dtype not in (np.float64, np.float32) or Y.size != 2 * fold or not Y.flags["C_CONTIGUOUS"]:
        raise ValueError("Y must be a contiguous array of 2*fold float64 or float32 elements")
    L = lib()
    if Y.dtype == np.float64:
        f = L.binned_dmdrescale
        f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_int]
    else:
        f = L.binned_smsrescale
        f.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_int]
    f.restype = None
    f(fold, new_scale, scale, Y.ctypes.data, 1, Y.ctypes.data + Y.itemsize * fold, 1)
    return Y


def allreduce_binned_ssq(Y, scale, fold, group=None):
    """Scaled binned sums of squares (the value and scale binnedBLAS_dbdssq /
    sbsssq return) merged over ranks: the scales all-reduced with MAX, every
    value rescaled to it here, then the values exchanged as allreduce_binned
    does. Returns the common scale (binned_dbdbaddsq returns the larger one).
    Bit-identical to merging the values with binned_dbdbaddsq in any order
    where those rescales are exact, i.e. unless a rescale makes a primary
    subnormal (there the reference's own result depends on its merge order).
    Collective."""
    import torch
    import torch.distributed as dist

    ts = torch.tensor([float(scale)], dtype=torch.float64)
    dist.all_reduce(ts, op=dist.ReduceOp.MAX, group=group)
    S = Y.dtype.type(ts.item())
    if S != Y.dtype.type(scale):
        rescale(Y, Y.dtype.type(scale), S, fold)
    allreduce_binned(Y, fold, group)
    return S
