"""Shared restatements for the edit side-channel tests: numpy bit packing, the
MPI row-major rank order of the blocks, synthetic (f, g0, g) triples with every
kind of voxel, and the expected symbols and clamp values per block."""
import numpy as np


def pack_numpy(sym, nbits):
    """LSB-first bit stream of the symbols, zero-padded to whole bytes."""
    bits = ((sym.astype(np.uint16)[:, None] >> np.arange(nbits)) & 1).astype(np.uint8)
    return np.packbits(bits.ravel(), bitorder="little")


def rank_blocks(shape, sub):
    """Core slices of the rank blocks: MPI_Cart_create(reorder = 0) numbers ranks
    row-major over dims {PX, PY, PZ}, so rank = (cx * PY + cy) * PZ + cz."""
    nz, ny, nx = shape
    px, py, pz = sub

    def split(n, parts, c):
        b, r = divmod(n, parts)
        ln = b + (1 if c < r else 0)
        return slice(c * b + min(c, r), c * b + min(c, r) + ln)

    out = []
    for r in range(px * py * pz):
        cx, cy, cz = r // (py * pz), (r // pz) % py, r % pz
        out.append((split(nz, pz, cz), split(ny, py, cy), split(nx, px, cx)))
    return out


def synthetic(shape, dtype, eb, q, seed, zero_clamps=True):
    """f, g0, g with every kind of voxel: untouched, lowered by k deltas,
    clamped (f - eb), clamped to exactly 0 (f == eb), and raised (g > g0)."""
    rng = np.random.default_rng(seed)
    T = np.dtype(dtype).type
    f = rng.uniform(-1, 1, shape).astype(dtype)
    g0 = (f + rng.uniform(-0.5, 0.5, shape).astype(dtype) * T(eb)).astype(dtype)
    delta = T(eb / q)
    kind = rng.integers(0, 5, shape)
    k = rng.integers(1, q, shape) if q > 1 else np.ones(shape, np.int64)
    g = g0.copy()
    low = kind == 1
    g[low] = (g0[low] - (k[low].astype(dtype) * delta)).astype(dtype)
    clamp = kind == 2
    g[clamp] = (f[clamp] - T(eb)).astype(dtype)
    if zero_clamps:
        zc = kind == 3
        f[zc] = T(eb)
        g0[zc] = T(eb)
        g[zc] = (f[zc] - T(eb)).astype(dtype)
    up = kind == 4
    g[up] = (g0[up] + delta).astype(dtype)
    return f, g0, g


def expected(f, g0, g, eb, q, sub, dtype):
    """Independent restatement in numpy of the symbols and kept clamp values per
    rank block (reference clamp mode: the clamp value is f - eb in T)."""
    T = np.dtype(dtype).type
    delta = T(eb / q)
    lowered = (g != g0) & (g < g0)
    cval = (f - T(eb)).astype(dtype)
    is_clamp = lowered & (g == cval)
    with np.errstate(invalid="ignore", over="ignore"):
        r = ((g0 - g) / delta).astype(dtype)
    rounded = np.where(r >= 0, np.floor(r.astype(np.float64) + 0.5), 0)
    sym = np.where(lowered, np.where(is_clamp, q, rounded), 0).astype(np.int64)
    out = []
    for sl in rank_blocks(f.shape, sub):
        s = sym[sl].ravel()
        keep = (is_clamp & (cval != 0))[sl].ravel()
        out.append((s, cval[sl].ravel()[keep]))
    return out, int(lowered.sum())
