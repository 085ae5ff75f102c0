"""GPU parity of the batch entry point (pmsz_correct_batch): every field of a
batch equals the oracle's correction of that field alone (field bytes, delta
counters, statistics) and the plain GPU correction, for mixed shapes, bounds,
q and clamp modes, both loop modes and the sweep caps."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STAT_KEYS = ("edit_sweeps", "check_sweeps", "outer_iters", "ite", "total_targets",
             "clamped_voxels", "touched_voxels", "status", "final_cnt")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    return torch.device("cuda:0")


def _fields(n, seed, dtype=np.float32, lo=8, hi=40):
    from pmsz_anonymous_amd import fieldio
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        shape = tuple(int(v) for v in rng.integers(lo, hi, size=3))
        eb_rel = float(10 ** rng.uniform(-3.5, -2))
        comp = "sz3" if k % 2 == 0 or not fieldio.zfp_available() else "zfp"
        out.append(fieldio.perlin_case(shape, 100 + seed + k, 4 + k % 3, 3, eb_rel, comp, dtype))
    return out


def _run_batch(cases, dev, per_field=None, **kw):
    from pmsz_anonymous_amd import engine
    fs = [torch.from_numpy(f).to(dev) for f, _, _ in cases]
    gs = [torch.from_numpy(g).to(dev) for _, g, _ in cases]
    dcs = [torch.zeros(f.shape, dtype=torch.uint8, device=dev) for f, _, _ in cases]
    sts = engine.correct_batch(fs, gs, [eb for _, _, eb in cases], dcs=dcs, per_field=per_field,
                               raise_not_converged=False, **kw)
    return [g.cpu().numpy() for g in gs], [d.cpu().numpy() for d in dcs], sts


def _check(cases, dev, per_field=None, **kw):
    print(f"batch of {len(cases)} kw={kw} per_field={per_field}", flush=True)
    gg, dd, sts = _run_batch(cases, dev, per_field=per_field, **kw)
    for k, (f, g0, eb) in enumerate(cases):
        extra = dict(kw)
        if per_field:
            extra.update(per_field[k])
        go, dco, sto, _, _ = O.correct(f, g0, eb, **extra)
        assert gg[k].tobytes() == go.tobytes(), f"field {k} differs"
        assert np.array_equal(dd[k], dco), f"field {k} counters differ"
        sto = dict(sto)
        sto["touched_voxels"] = int((dco > 0).sum())
        for key in STAT_KEYS:
            assert sts[k][key] == sto[key], (k, key, sts[k][key], sto[key])
    return gg, sts


@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
def test_mixed_batch_f32(dev, mode):
    cases = _fields(7, 1)
    per = [{"q": 5 if k % 3 else 3, "clamp_mode": k % 2} for k in range(len(cases))]
    _check(cases, dev, per_field=per, mode=mode)


def test_batch_f64(dev):
    _check(_fields(5, 2, dtype=np.float64), dev)


@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
@pytest.mark.parametrize("cap", [1, 2, 4])
def test_batch_caps(dev, mode, cap):
    _check(_fields(4, 3, lo=12, hi=30), dev, mode=mode, max_sweeps=cap)


def test_batch_equals_plain(dev):
    from pmsz_anonymous_amd import engine
    cases = _fields(6, 4, lo=16, hi=48)
    gg, sts = _run_batch(cases, dev)[0::2]
    for k, (f, g0, eb) in enumerate(cases):
        tg = torch.from_numpy(g0).to(dev)
        _, st = engine.correct(torch.from_numpy(f).to(dev), tg, eb)
        assert tg.cpu().numpy().tobytes() == gg[k].tobytes()
        for key in ("edit_sweeps", "total_targets", "touched_voxels", "clamped_voxels"):
            assert st[key] == sts[k][key], (k, key)


def test_batch_repeat_deterministic(dev):
    cases = _fields(5, 5)
    a = _run_batch(cases, dev)
    b = _run_batch(cases, dev)
    for k in range(len(cases)):
        assert a[0][k].tobytes() == b[0][k].tobytes() and np.array_equal(a[1][k], b[1][k])


def test_batch_rejects_mixed_modes(dev):
    from pmsz_anonymous_amd import engine
    cases = _fields(2, 6)
    fs = [torch.from_numpy(f).to(dev) for f, _, _ in cases]
    gs = [torch.from_numpy(g).to(dev) for _, g, _ in cases]
    with pytest.raises(engine.PmszError) as ei:
        engine.correct_batch(fs, gs, [e for _, _, e in cases],
                             per_field=[{"mode": O.LAZY}, {"mode": O.SYNC}])
    assert ei.value.code == engine.EINVAL
