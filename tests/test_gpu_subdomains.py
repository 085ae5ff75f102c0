"""GPU parity of the subdomain path (pmsz_correct with sub != 1,1,1) against the
oracle's literal padded-block restatement: corrected field bytes, delta counters
and the run statistics, bit-exact, in both loop modes and both ghost policies,
uneven and thin splits, both clamp modes, and the sweep caps."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STAT_KEYS = ("edit_sweeps", "check_sweeps", "outer_iters", "ite", "total_targets",
             "clamped_voxels", "status", "final_cnt")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    return torch.device("cuda:0")


def _case(shape, seed, dtype=np.float32, eb_rel=1e-3, freq=5, oct_=4, comp="sz3"):
    from pmsz_anonymous_amd import fieldio
    return fieldio.perlin_case(shape, seed, freq, oct_, eb_rel, comp, dtype)


def _compare(f, g0, eb, dev, **kw):
    from pmsz_anonymous_amd import engine
    print(f"case shape={f.shape} dtype={f.dtype} eb={eb:.3e} kw={kw}", flush=True)
    go, dco, sto, _, _ = O.correct(f, g0, eb, **kw)
    tf = torch.from_numpy(f).to(dev)
    tg = torch.from_numpy(g0).to(dev)
    dc = torch.zeros(f.shape, dtype=torch.uint8, device=dev)
    _, stg = engine.correct(tf, tg, eb, dc=dc, raise_not_converged=False, **kw)
    gg = tg.cpu().numpy()
    assert gg.tobytes() == go.tobytes(), "corrected field differs"
    assert np.array_equal(dc.cpu().numpy(), dco), "delta counters differ"
    for k in STAT_KEYS:
        assert stg[k] == sto[k], (k, stg[k], sto[k])
    assert stg["touched_voxels"] == int((dco > 0).sum())
    return go, sto


@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
@pytest.mark.parametrize("policy", [O.GHOST_REFERENCE, O.GHOST_FULL])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_2x2x2_parity(dev, mode, policy, dtype):
    f, g0, eb = _case((40, 44, 48), 31, dtype=dtype)
    _compare(f, g0, eb, dev, sub=(2, 2, 2), mode=mode, ghost_policy=policy)


@pytest.mark.parametrize("sub", [(2, 1, 1), (1, 3, 1), (1, 1, 4), (3, 2, 2), (4, 3, 2)])
@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
def test_uneven_splits(dev, sub, mode):
    # extents not divisible by the split: the first n % parts blocks are larger
    f, g0, eb = _case((29, 35, 41), 32, eb_rel=3e-3)
    _compare(f, g0, eb, dev, sub=sub, mode=mode, ghost_policy=O.GHOST_REFERENCE)


@pytest.mark.parametrize("sub", [(4, 1, 1), (1, 1, 5), (3, 3, 3)])
def test_thin_blocks(dev, sub):
    # cores of 1..3 cells along an axis: the 2-layer shell covers the whole block
    f, g0, eb = _case((6, 7, 9), 33, eb_rel=1e-2)
    _compare(f, g0, eb, dev, sub=sub, ghost_policy=O.GHOST_FULL)


@pytest.mark.parametrize("clamp", [O.CLAMP_REFERENCE, O.CLAMP_BOUND_SAFE])
@pytest.mark.parametrize("q", [2, 5])
def test_clamp_and_q(dev, clamp, q):
    f, g0, eb = _case((30, 32, 34), 34, eb_rel=1e-2, freq=7)
    _compare(f, g0, eb, dev, sub=(2, 2, 2), clamp_mode=clamp, q=q)


@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
@pytest.mark.parametrize("cap", [1, 2, 3, 5])
def test_sweep_caps(dev, mode, cap):
    f, g0, eb = _case((24, 26, 28), 35, eb_rel=3e-3, freq=6)
    _compare(f, g0, eb, dev, sub=(2, 2, 1), mode=mode, max_sweeps=cap)


def test_single_split_matches_plain(dev):
    # 1x1x1 is the plain path; a 2x2x2 full-policy run converges to a field whose
    # extrema and labels equal those of f, like the plain run
    from pmsz_anonymous_amd import engine
    f, g0, eb = _case((32, 32, 32), 36)
    g_sub, st = _compare(f, g0, eb, dev, sub=(2, 2, 2), ghost_policy=O.GHOST_FULL)
    assert st["final_cnt"] == [0, 0, 0]
    a0, d0 = O.labels(f)
    a1, d1 = O.labels(g_sub)
    assert np.array_equal(a0, a1) and np.array_equal(d0, d1)
    tf = torch.from_numpy(f).to(dev)
    assert torch.equal(engine.codes(tf), engine.codes(torch.from_numpy(g_sub).to(dev)))


def test_subdomain_rejects_nan(dev):
    from pmsz_anonymous_amd import engine
    f, g0, eb = _case((16, 16, 16), 37)
    g1 = g0.copy()
    g1[9, 3, 12] = np.nan
    with pytest.raises(engine.PmszError) as ei:
        engine.correct(torch.from_numpy(f).to(dev), torch.from_numpy(g1).to(dev), eb,
                       sub=(2, 2, 2))
    assert ei.value.code == engine.EINVAL


@pytest.mark.slow
@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
def test_c3_shape_properties(dev, mode):
    # C3's layout (2x2x2 subdomains, f64, eb = 1e-4 x range) at 192^3: converged,
    # every voxel within the bound, only lowered, labels of g equal those of f
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((192, 192, 192), 2, 16, 6, 1e-4, "sz3", np.float64)
    tf = torch.from_numpy(f).to(dev)
    tg = torch.from_numpy(g0).to(dev)
    _, st = engine.correct(tf, tg, eb, sub=(2, 2, 2), mode=mode,
                           ghost_policy=O.GHOST_FULL)
    assert st["final_cnt"] == [0, 0, 0]
    g = tg.cpu().numpy()
    assert fieldio.bound_violations(f, g, eb)[1] == 0
    assert (g <= g0).all()
    a0, d0 = engine.labels(tf)
    a1, d1 = engine.labels(tg)
    assert torch.equal(a0, a1) and torch.equal(d0, d1)
