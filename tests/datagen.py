"""Seeded synthetic inputs shared by the tests and the benchmark.

All draws use numpy's PCG64, which is platform independent, so the same seed
gives the same vector here and on the GPU host.
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def uniform(n, seed, lo=-1.0, hi=1.0):
    return rng(seed).uniform(lo, hi, n)


def wide(n, seed, emin=-40, emax=40):
    """x = s * m * 2^e, s = +-1, m ~ U[1,2), e ~ U{emin..emax}."""
    g = rng(seed)
    s = g.choice(np.array([-1.0, 1.0]), n)
    m = g.uniform(1.0, 2.0, n)
    e = g.integers(emin, emax + 1, n)
    return s * np.ldexp(m, e)


def ill_conditioned(n, seed, noise=1e-12):
    """N/2 pairs (x, -x + noise*u) interleaved, then shuffled by `seed`.

    The multiset of values depends only on n (it is drawn from seed 0);
    `seed` only permutes it, so a reproducible sum must not depend on it."""
    g = rng(0)
    half = n // 2
    x = g.uniform(-1.0, 1.0, half) * np.exp2(g.integers(-20, 21, half).astype(float))
    v = np.empty(n)
    v[0:2 * half:2] = x
    v[1:2 * half:2] = -x + noise * g.uniform(-1.0, 1.0, half)
    if n > 2 * half:
        v[-1] = noise * g.uniform(-1.0, 1.0)
    if seed == 0:
        return v
    return rng(seed).permutation(v)


def normal(shape, seed, scale=1.0):
    return rng(seed).normal(0.0, scale, shape)


def special(kind, n, seed):
    """Inputs with non-finite or extreme members."""
    x = uniform(n, seed)
    g = rng(seed + 1000)
    if kind == "nan":
        x[g.integers(0, n)] = np.nan
    elif kind == "pinf":
        x[g.integers(0, n)] = np.inf
    elif kind == "ninf":
        x[g.integers(0, n)] = -np.inf
    elif kind == "both_inf":
        i, j = g.choice(n, 2, replace=False)
        x[i], x[j] = np.inf, -np.inf
    elif kind == "inf_nan":
        i, j = g.choice(n, 2, replace=False)
        x[i], x[j] = np.inf, np.nan
    elif kind == "big":  # index-0 (compressed top bin) path
        x = x * np.ldexp(1.0, 1020)
        x[g.integers(0, n)] = 1.7976931348623157e308
    elif kind == "tiny":  # subnormals and near-underflow
        x = x * np.ldexp(1.0, -1060)
    elif kind == "mixed_range":
        x = x * np.exp2(g.integers(-1000, 1000, n).astype(float))
    elif kind == "zeros":
        x = np.zeros(n)
    elif kind == "neg_zeros":
        x = -np.zeros(n)
    elif kind == "constant":
        x = np.full(n, 0.1)
    else:
        raise ValueError(kind)
    return x


# ---------------------------------------------------------- single precision --
def wide_s(n, seed, emin=-20, emax=20):
    """float32 x = s * m * 2^e, s = +-1, m ~ U[1,2) (rounded to float), e ~ U{emin..emax}."""
    return wide(n, seed, emin, emax).astype(np.float32)


def special_s(kind, n, seed):
    """float32 inputs with non-finite or extreme members."""
    if kind == "big":  # index-0 (compressed top bin) path of float_binned
        x = (uniform(n, seed) * np.ldexp(1.0, 125)).astype(np.float32)
        x[rng(seed + 1000).integers(0, n)] = np.float32(3.4028234663852886e38)
        return x
    if kind == "tiny":  # float subnormals and near-underflow
        return (uniform(n, seed) * np.ldexp(1.0, -135)).astype(np.float32)
    if kind == "mixed_range":
        g = rng(seed + 1000)
        return (uniform(n, seed) * np.exp2(g.integers(-140, 120, n).astype(float))).astype(np.float32)
    return special(kind, n, seed).astype(np.float32)


# ---------------------------------------------- infinity-path corner cases --
INF_CORNERS = ["neg_overflow_then_pinf", "pinf_then_neg_big", "pos_overflow_then_pinf", "top_state_then_ninf",
               "top_state_small_then_ninf", "overflow_at_block_edge", "two_inf_blocks", "late_block", "fuzz"]


def inf_corner(kind, n, seed, op="sum", dtype=np.float64, block=None):
    """Inputs that reach the reference's infinity path (dmdsum.ccog:66-78,
    dmddot.ccog:68-81): at its first block holding an infinite term the
    reference adds that block's terms one by one to its top primary, so a
    sequential overflow before the infinity decides between NaN and the
    infinity. `block` is the reference's block length for (op, dtype); terms
    are x (sum) or x*y (dot). Returns (x, y)."""
    f = np.finfo(dtype)
    big = float(f.max) * 0.95  # two of them overflow
    blk = block or ((1024 if op == "dot" else 2048) if dtype == np.float64 else 512)
    g = rng(seed)
    x = uniform(n, seed).astype(dtype)
    y = np.ones(n, dtype) if op == "dot" else None
    b = min(max(1, int(g.integers(1, 4))), max(0, n // blk - 1))  # the block that gets the infinity
    p = b * blk + int(g.integers(0, blk // 2))
    if kind == "neg_overflow_then_pinf":
        x[p], x[p + 1], x[p + 5] = -big, -big, np.inf
    elif kind == "pinf_then_neg_big":
        x[p], x[p + 1], x[p + 2] = np.inf, -big, -big
    elif kind == "pos_overflow_then_pinf":
        x[p], x[p + 1], x[p + 3] = big, big, np.inf
    elif kind == "top_state_then_ninf":
        # an earlier block raises the state to index 0; then one big term
        # overflows the top primary before the infinity
        x[p - blk] = big
        x[p], x[p + 2] = float(f.max) * 0.6, -np.inf
    elif kind == "top_state_small_then_ninf":
        x[p - blk] = big
        x[p], x[p + 2] = 1e-3 * float(f.max), -np.inf
    elif kind == "overflow_at_block_edge":
        e = b * blk
        x[e - 2], x[e - 1], x[e] = -big, -big, np.inf
    elif kind == "two_inf_blocks":
        x[p], x[p + 1], x[p + 4] = -big, -big, np.inf
        x[min(n - 1, p + blk + 7)] = np.inf
    elif kind == "late_block":
        q = (n // blk - 1) * blk + 3
        x[q], x[q + 1], x[q + 2] = -big, -big, np.inf
    elif kind == "fuzz":
        k = int(g.integers(2, 12))
        pos = g.choice(n, k, replace=False)
        x[pos] = g.choice(np.array([-1.0, 1.0]), k) * g.uniform(0.3, 0.99, k) * float(f.max)
        x[int(g.integers(0, n))] = np.inf if g.integers(0, 2) else -np.inf
    else:
        raise ValueError(kind)
    if op == "dot":
        # the same terms as products: x*y with y = 1 except a few exact
        # splittings; for odd seeds every infinite term is a product of two
        # finite operands that overflows on its own
        sp = g.choice(n, 3, replace=False)
        for i in sp:
            if np.isfinite(x[i]) and abs(x[i]) < 1e-3 * float(f.max):
                x[i], y[i] = x[i] * 2.0, dtype(0.5)
        if seed % 2:
            r = dtype(np.sqrt(float(f.max)) * 1.5)
            for i in np.nonzero(np.isinf(x))[0]:
                x[i], y[i] = (r if x[i] > 0 else -r), r
    return x, y
