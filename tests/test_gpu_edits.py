"""GPU parity of the edit side channel (pmsz_encode_edits) against the oracle,
bit-exact: packed symbol streams, kept clamp values, per-block layout and the
edited count; corrected fields and synthetic triples with every kind of voxel,
one field and rank blocks, both value types, q from 1 to 255, unaligned
pointers, many tiles; and at the full 512^3 size, properties the encoding fixes."""
import numpy as np
import pytest

import pyoracle as O
from editcheck import rank_blocks, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    return torch.device("cuda:0")


def _tensors(dev, *arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrays]


def _compare(f, g0, g, eb, dev, tensors=None, **kw):
    from pmsz_anonymous_amd import engine
    print(f"encode shape={f.shape} dtype={f.dtype} kw={kw}", flush=True)
    opk, ovals, obl, oed, orc = O.encode_edits(f, g0, g, eb, **kw)
    assert orc == 0
    tf, tg0, tg = tensors if tensors is not None else _tensors(dev, f, g0, g)
    pk, vals, bl, info = engine.encode_edits(tf, tg0, tg, eb, **kw)
    assert bl == obl
    assert info["edited"] == oed
    assert info["values"] == ovals.size
    assert info["nbits"] == O.bits_for_max(kw.get("q", 5))
    pk = pk.cpu().numpy()
    for b in bl:
        lo, hi = b["packed_offset"], b["packed_offset"] + b["packed_bytes"]
        assert pk[lo:hi].tobytes() == opk[lo:hi].tobytes(), b
    assert vals.cpu().numpy().tobytes() == ovals.tobytes()
    return bl, info


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sub", [(1, 1, 1), (2, 1, 3), (2, 2, 2)])
def test_encode_corrected_field(dev, dtype, sub):
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((40, 44, 52), 5, 5, 4, 3e-3, "sz3", dtype)
    tf, tg = _tensors(dev, f, g0)
    engine.correct(tf, tg, eb)
    _compare(f, g0, tg.cpu().numpy(), eb, dev, sub=sub)


@pytest.mark.parametrize("q", [1, 2, 5, 17, 200, 255])
@pytest.mark.parametrize("clamp", [O.CLAMP_REFERENCE, O.CLAMP_BOUND_SAFE])
def test_encode_synthetic_q(dev, q, clamp):
    f, g0, g = synthetic((23, 29, 31), np.float32, 1e-2, q, seed=q)
    _compare(f, g0, g, 1e-2, dev, q=q, clamp_mode=clamp, sub=(1, 2, 1))
    _compare(f, g0, g, 1e-2, dev, q=q, clamp_mode=clamp)


@pytest.mark.parametrize("shape", [(1, 1, 2), (1, 3, 5), (2, 1, 2047), (1, 1, 2049),
                                   (3, 7, 1000)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_encode_ragged(dev, shape, dtype):
    f, g0, g = synthetic(shape, dtype, 1e-2, 5, seed=sum(shape))
    _compare(f, g0, g, 1e-2, dev)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_encode_many_tiles(dev, dtype):
    # about 1000 tiles of 2048 voxels, a third of them clamp voxels: the
    # cross-tile scan places every value
    f, g0, g = synthetic((130, 128, 128), dtype, 1e-2, 5, seed=2)
    _compare(f, g0, g, 1e-2, dev, clamp_mode=O.CLAMP_REFERENCE)
    _compare(f, g0, g, 1e-2, dev, clamp_mode=O.CLAMP_REFERENCE, sub=(3, 1, 2))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_encode_unaligned(dev, dtype):
    f, g0, g = synthetic((17, 19, 21), dtype, 1e-2, 5, seed=4)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    ts = []
    for a in (f, g0, g):
        buf = torch.zeros(a.size + 1, dtype=tdt, device=dev)
        t = buf[1:].view(a.shape)
        t.copy_(torch.from_numpy(a))
        assert t.data_ptr() % 16 != 0
        ts.append(t)
    _compare(f, g0, g, 1e-2, dev, tensors=ts)


def test_encode_symbol_out_of_range(dev):
    from pmsz_anonymous_amd import engine
    f = np.ones((4, 4, 4), np.float32)
    g0 = f.copy()
    g = f.copy()
    g[1, 1, 1] = g0[1, 1, 1] - np.float32(9 * (1e-2 / 5))
    tf, tg0, tg = _tensors(dev, f, g0, g)
    with pytest.raises(engine.PmszError) as ei:
        engine.encode_edits(tf, tg0, tg, 1e-2)
    assert ei.value.code == engine.EINVAL


def test_encode_then_files_round_trip(dev, tmp_path):
    from pmsz_anonymous_amd import engine, fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    f, g0, eb = fieldio.perlin_case((33, 34, 35), 2, 5, 4, 3e-3, "sz3", np.float32)
    tf, tg = _tensors(dev, f, g0)
    tg0 = tg.clone()
    engine.correct(tf, tg, eb)
    pk, vals, bl, info = engine.encode_edits(tf, tg0, tg, eb, sub=(2, 1, 2))
    prefix = str(tmp_path / "e")
    fieldio.write_edits(prefix, np.float32, bl, pk.cpu().numpy(), vals.cpu().numpy())
    sym, v2 = fieldio.read_edits(prefix, np.float32, 5, [b["voxels"] for b in bl], f.size)
    want, _ = engine.edit_symbols(tf, tg0, tg, eb)
    want = want.cpu().numpy()
    assert sym.tobytes() == np.concatenate(
        [want[sl].ravel() for sl in rank_blocks(f.shape, (2, 1, 2))]).tobytes()
    assert v2.tobytes() == vals.cpu().numpy().tobytes()


def _unpack(pk, n, nbits):
    """LSB-first nbits-bit symbols of a packed byte tensor, on the device."""
    bits = ((pk.to(torch.int32)[:, None] >> torch.arange(8, device=pk.device)) & 1)
    bits = bits.reshape(-1)[: n * nbits].reshape(n, nbits)
    w = (1 << torch.arange(nbits, device=pk.device)).to(torch.int32)
    return (bits * w).sum(1).to(torch.uint8)


@pytest.mark.slow
def test_encode_512cube_properties(dev):
    # full size: the packed stream unpacks to pmsz_edit_symbols' symbols, and the
    # values are g at the voxels lowered to exactly f - eb, zeros left out
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((512, 512, 512), 1, 8, 5, 1e-3, "sz3", np.float32)
    tf, tg = _tensors(dev, f, g0)
    del f, g0
    tg0 = tg.clone()
    engine.correct(tf, tg, eb, clamp_mode=O.CLAMP_REFERENCE)
    pk, vals, bl, info = engine.encode_edits(tf, tg0, tg, eb, clamp_mode=O.CLAMP_REFERENCE)
    sym, ed = engine.edit_symbols(tf, tg0, tg, eb, clamp_mode=O.CLAMP_REFERENCE)
    assert info["edited"] == ed > 0
    n = tf.numel()
    assert bl[0]["packed_bytes"] == (n * 3 + 7) // 8
    assert torch.equal(_unpack(pk, n, 3), sym.reshape(-1))
    c = tf - torch.tensor(eb, dtype=torch.float32)
    keep = (tg < tg0) & (tg == c) & (c != 0)
    assert torch.equal(vals, tg[keep])
