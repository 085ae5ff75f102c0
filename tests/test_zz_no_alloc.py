"""pmsz_encode_edits_noalloc, the encoder entry point that allocates nothing
(its scratch is the library's fixed per-device array), against pmsz_encode_edits
on the same inputs, bit for bit: packed streams, clamp values, block layout and
totals.  Two calls in a row, the second on a smaller field (the array keeps the
first call's contents), and requests beyond the array's limits rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    return torch.device("cuda:0")


def _corrected(dev, shape, seed, dtype, eb_rel):
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case(shape, seed, 5, 4, eb_rel, "sz3", dtype)
    tf = torch.from_numpy(f).to(dev)
    tg0 = torch.from_numpy(g0).to(dev)
    tg = tg0.clone()
    engine.correct(tf, tg, eb)
    return tf, tg0, tg, eb


def _same(dev, tf, tg0, tg, eb, **kw):
    from pmsz_anonymous_amd import engine
    print(f"noalloc encode shape={tuple(tf.shape)} dtype={tf.dtype} kw={kw}", flush=True)
    pk, vals, bl, info = engine.encode_edits(tf, tg0, tg, eb, **kw)
    pk = pk.cpu().numpy().copy()
    vals = vals.cpu().numpy().copy()
    # outputs prefilled with garbage: every byte the entry point reports must be written
    npk = torch.full((max(pk.size, 16),), 0xA5, dtype=torch.uint8, device=dev)
    nvals = torch.full((tf.numel(),), 7.0, dtype=tf.dtype, device=dev)
    npk, nvals, nbl, ninfo = engine.encode_edits_noalloc(tf, tg0, tg, eb, packed=npk, vals=nvals,
                                                         **kw)
    torch.cuda.synchronize()
    assert nbl == bl and ninfo == info
    npk = npk.cpu().numpy()
    for b in bl:
        lo, hi = b["packed_offset"], b["packed_offset"] + b["packed_bytes"]
        assert npk[lo:hi].tobytes() == pk[lo:hi].tobytes(), b
    assert nvals.cpu().numpy().tobytes() == vals.tobytes()
    assert info["edited"] > 0 and info["values"] > 0
    return info


def test_noalloc_matches_encode_edits_twice(dev):
    # first call: a larger float field in one block; second: a smaller double field
    # in 2 x 1 x 3 rank blocks
    a = _corrected(dev, (72, 80, 96), 11, np.float32, 2e-3)
    _same(dev, *a)
    b = _corrected(dev, (40, 44, 52), 5, np.float64, 3e-3)
    _same(dev, *b, sub=(2, 1, 3))
    # and the first request again after the smaller one
    _same(dev, *a)


def test_noalloc_rejects_beyond_limits(dev):
    from pmsz_anonymous_amd import engine
    L = engine.lib()
    dummy = ctypes.c_void_p(1 << 20)
    for shape, sub in (((1200, 1200, 1200), (1, 1, 1)),   # tiles beyond the 4 MiB array
                       ((64, 64, 64), (8, 8, 8))):        # more than 256 rank blocks
        p = engine.make_params(shape, "float32", 1e-3, sub=sub)
        nb = sub[0] * sub[1] * sub[2]
        blocks = (engine.EditBlock * nb)()
        info = engine.EditInfo()
        rc = L.pmsz_encode_edits_noalloc(ctypes.byref(p), dummy, dummy, dummy, dummy, dummy,
                                         blocks, ctypes.byref(info), None)
        assert rc == engine.ETOOBIG, (shape, sub, rc)
