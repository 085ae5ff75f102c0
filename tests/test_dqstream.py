"""The codec's compressed stream in the oracle (no GPU), pinned by an
independent decoder written from the stream's layout (tests/dqstream_ref.py):
its g0 equals the codec's g0 bit for bit, over value types, rank grids, escapes
and verbatim voxels; the code of every block is a complete prefix code
(Kraft sum 1) as short as a Huffman code built with heapq (when no length limit
bites), and the stream is within a stated bound of the entropy estimate."""
import heapq
import math
import struct

import numpy as np
import pytest

import pyoracle as O
import dqstream_ref as R


def _case(shape, seed, dtype, rel=1e-3):
    from pmsz_anonymous_amd import fieldio
    f = fieldio.perlin(shape, seed, 4, 3, dtype=dtype)
    lo, hi = fieldio.value_range(f)
    return f, rel * (hi - lo)


def _lengths(stream):
    b = bytes(stream)
    nb = struct.unpack_from("<I", b, 44)[0]
    o = 64 + 64 * nb
    return [np.frombuffer(b, np.uint8, 1025, o + 1025 * r) for r in range(nb)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sub", [(1, 1, 1), (2, 1, 3), (2, 2, 2)])
def test_independent_decoder_recovers_g0(dtype, sub):
    f, eb = _case((21, 26, 30), 3, dtype)
    g0, stream = O.dq_compress(f, eb, sub)
    wg0, _ = O.codec_dq(f, eb, sub)
    assert g0.tobytes() == wg0.tobytes()
    rg0, _ = R.decode(stream)
    assert rg0.tobytes() == g0.tobytes()
    assert O.dq_decompress(stream).tobytes() == g0.tobytes()


def test_escapes_verbatim_and_limits():
    rng = np.random.default_rng(4)
    f = (rng.standard_normal((9, 10, 11)) * 500).astype(np.float64)
    f[2, 3, 4] = 1e12          # beyond the quantum range: k = 0, verbatim
    f[3, 3, 3] = 2.0 ** 30 - 0.25  # rounds to 2^30 at eb = 0.5: verbatim
    f[4, 4, 4] = np.nan        # verbatim
    g0, stream = O.dq_compress(f, 0.5)
    _, h = O.codec_dq(f, 0.5)
    rg0, (nesc, nverb, _) = R.decode(stream)
    assert nesc == int(h[0, O.DQ_BINS]) > 0 and nverb == int(h[0, O.DQ_BINS + 1]) == 3
    assert rg0.tobytes() == g0.tobytes()
    assert np.isnan(rg0[4, 4, 4]) and rg0[2, 3, 4] == 1e12


def test_constant_field_one_symbol_per_block():
    f = np.full((12, 13, 14), 7 * 0.5, np.float32)   # quantum 7 everywhere
    g0, stream = O.dq_compress(f, 0.25, (2, 1, 1))
    rg0, _ = R.decode(stream)
    assert rg0.tobytes() == g0.tobytes() == f.tobytes()


def _heap_cost(counts):
    h = [c for c in counts if c]
    if len(h) == 1:
        return h[0]
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a, b = heapq.heappop(h), heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


@pytest.mark.parametrize("rel", [1e-2, 1e-4])
def test_codes_complete_and_optimal(rel):
    f, eb = _case((24, 25, 26), 7, np.float32, rel)
    g0, stream = O.dq_compress(f, eb, (2, 1, 1))
    _, hist = O.codec_dq(f, eb, (2, 1, 1))
    for lens, h in zip(_lengths(stream), hist):
        counts = [int(c) for c in h[:O.DQ_BINS + 1]]
        used = [c for c in counts if c]
        assert all((l > 0) == (c > 0) for l, c in zip(lens, counts))
        if len(used) > 1:
            assert sum(2.0 ** -int(l) for l in lens if l) == 1.0  # complete prefix code
        assert max(lens) <= 24
        assert sum(c * int(l) for c, l in zip(counts, lens)) == _heap_cost(counts)


def test_size_within_bound_of_entropy_estimate():
    # per block: entropy bits <= code bits < entropy bits + symbols (Huffman); the
    # stream adds the header, block records, code lengths, chunk offsets, the
    # word padding of each chunk, the escapes and the verbatim values
    f, eb = _case((64, 64, 64), 1, np.float32)
    _, stream = O.dq_compress(f, eb)
    _, hist = O.codec_dq(f, eb)
    n = f.size
    h = hist[0, :O.DQ_BINS + 1].astype(np.float64)
    ent = float((h[h > 0] * np.log2(n / h[h > 0])).sum())
    chunks = math.ceil(n / 2048)
    fixed = 64 + 64 + 1025 + 16 * chunks + 8 + 4 * int(hist[0, O.DQ_BINS]) + \
        12 * int(hist[0, O.DQ_BINS + 1])
    assert ent / 8 <= stream.size <= (ent + n) / 8 + fixed + 4 * chunks
    assert f.nbytes / stream.size > 4.0
