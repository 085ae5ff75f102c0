"""The dual-quantisation Lorenzo codec of the oracle (the GPU compressor front
end's stand-in for SZ3, pMSz.cu:1536-1624) pinned against an independent numpy
restatement (exact round-half-away quantisation, exact bound check in rationals,
zero-padded 7-term Lorenzo per rank block) and a closed form: a constant field
quantises exactly and leaves one nonzero residual per block, at its first voxel."""
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as O  # noqa: E402
from editcheck import rank_blocks  # noqa: E402

R = O.DQ_RADIUS


def quantise(f, eb):
    """k = round half away from zero of f * (1 / (2 eb)), exactly; 0 where the
    rounded value reaches 2^30 in magnitude, and for NaN."""
    r = f.astype(np.float64) * (1.0 / (2.0 * eb))
    with np.errstate(invalid="ignore"):
        a = np.abs(r)
        fl = np.floor(a)
        k = (fl + (a - fl >= 0.5)) * np.sign(r)
        ok = np.abs(k) < 2.0 ** 30
    return np.where(ok, np.nan_to_num(k), 0).astype(np.int64)


def within(a, b, e):
    """|a - b| <= e in rationals; False when either value is NaN."""
    if a != a or b != b:
        return False
    return abs(Fraction(float(a)) - Fraction(float(b))) <= e


def dq_numpy(f, eb, sub):
    k = quantise(f, eb)
    v = (k.astype(np.float64) * (2.0 * eb)).astype(f.dtype)
    e = Fraction(float(eb))
    keep = np.array([within(a, b, e) for a, b in zip(f.ravel(), v.ravel())]).reshape(f.shape)
    g0 = np.where(keep, v, f).astype(f.dtype)
    hists = []
    for sl in rank_blocks(f.shape, sub):
        kb = np.pad(k[sl], ((1, 0), (1, 0), (1, 0)))
        c = kb[1:, 1:, 1:]
        pred = (kb[1:, 1:, :-1] + kb[1:, :-1, 1:] + kb[:-1, 1:, 1:] - kb[1:, :-1, :-1]
                - kb[:-1, 1:, :-1] - kb[:-1, :-1, 1:] + kb[:-1, :-1, :-1])
        d = (c - pred).ravel()
        h = np.zeros(O.DQ_BINS + 2, np.uint64)
        inr = np.abs(d) < R
        np.add.at(h, d[inr] + R, 1)
        h[O.DQ_BINS] = int((~inr).sum())
        h[O.DQ_BINS + 1] = int((~keep[sl]).sum())
        hists.append(h)
    return g0, np.array(hists)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sub", [(1, 1, 1), (2, 1, 3), (2, 2, 2)])
def test_codec_matches_numpy(dtype, sub):
    from pmsz_anonymous_amd import fieldio
    f = fieldio.perlin((6, 7, 9), 2, 4, 3, dtype=dtype)
    for eb in (1e-3, 3e-2):
        g0, h = O.codec_dq(f, eb, sub)
        wg0, wh = dq_numpy(f, eb, sub)
        assert g0.tobytes() == wg0.tobytes()
        assert np.array_equal(h, wh)


def test_codec_escapes_and_verbatim():
    # residuals beyond the radius, and values too large to quantise (k = 0, verbatim)
    rng = np.random.default_rng(1)
    f = (rng.standard_normal((5, 6, 7)) * 50).astype(np.float64)
    f[2, 3, 4] = 1e12
    g0, h = O.codec_dq(f, 1e-3)
    wg0, wh = dq_numpy(f, 1e-3, (1, 1, 1))
    assert g0.tobytes() == wg0.tobytes() and np.array_equal(h, wh)
    assert h[0, O.DQ_BINS] > 0 and h[0, O.DQ_BINS + 1] >= 1
    assert g0[2, 3, 4] == f[2, 3, 4]


def test_codec_quantum_limit_and_nan():
    # f / (2 eb) = 2^30 - 0.25 rounds to 2^30: k = 0 and the voxel is verbatim (2k + 1
    # would overflow an int32); 2^30 - 0.75 rounds to 2^30 - 1 and is kept; NaN is
    # verbatim and counted as such
    f = np.zeros((4, 5, 6), np.float64)
    f[1, 2, 3] = 2.0 ** 30 - 0.25
    f[2, 2, 2] = 2.0 ** 30 - 0.75
    f[3, 4, 5] = np.nan
    g0, h = O.codec_dq(f, 0.5)
    wg0, wh = dq_numpy(f, 0.5, (1, 1, 1))
    assert g0.tobytes() == wg0.tobytes() and np.array_equal(h, wh)
    assert g0[1, 2, 3] == f[1, 2, 3] and g0[2, 2, 2] == 2.0 ** 30 - 1
    assert np.isnan(g0[3, 4, 5])
    assert h[0, O.DQ_BINS + 1] == 2


@pytest.mark.parametrize("sub", [(1, 1, 1), (3, 2, 1)])
def test_codec_constant_field_closed_form(sub):
    eb = 0.25
    f = np.full((8, 9, 10), 7 * 2 * eb, np.float32)  # quantum 7 everywhere
    g0, h = O.codec_dq(f, eb, sub)
    assert g0.tobytes() == f.tobytes()
    for b, sl in enumerate(rank_blocks(f.shape, sub)):
        n = f[sl].size
        assert h[b, R] == n - 1 and h[b, R + 7] == 1
        assert h[b].sum() == n


def test_codec_bound_holds():
    from pmsz_anonymous_amd import fieldio
    import refcheck
    f = fieldio.perlin((10, 11, 12), 5, 5, 4, dtype=np.float32)
    g0, _ = O.codec_dq(f, 1e-4)
    assert refcheck.exact_within(f, g0, 1e-4)
