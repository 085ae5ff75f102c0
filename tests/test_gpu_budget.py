"""Step budgets against a simulation of the edits they grant (the test hook
pmsz_test_step_budget).  A budget B for a cell with counter d promises that the
edits made while the counter is below B are the plain step g - delta and honour
the bound, so the engine skips reading f for them.  For random cells -- values
from subnormal to large, of both signs and zero, bounds from 1e-9 of the value
up to the value, g anywhere within the bound and a little beyond it -- every
granted step is replayed with the value type's rounding and checked: bound-safe
mode exactly in rationals (|f - cand| <= eb), reference mode with the rounded
test of the reference (|fl(f - cand)| <= fl(eb), pMSz.cu:623-645)."""
from fractions import Fraction

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


def _group(rng, n, dt):
    """n cells of one launch: values around a base magnitude (both signs, some
    zeros), one bound relative to it, g within the bound and a little beyond."""
    lo, hi = (-40, 30) if dt == np.float32 else (-300, 200)
    m = 10.0 ** rng.uniform(lo, hi)
    f = (m * rng.uniform(-2, 2, n)).astype(dt)
    f[rng.random(n) < 0.1] = 0
    eb = max(m * 10.0 ** rng.uniform(-7, 0), float(np.finfo(dt).tiny))
    r = rng.uniform(-1.02, 1.02, n)
    r[rng.random(n) < 0.05] = 1.0
    r[rng.random(n) < 0.05] = -1.0
    g = (f.astype(np.float64) + r * eb).astype(dt)
    return f, g, eb


def _accepted(fv, cand, eb, mode, dt):
    if mode == 1:  # bound-safe: exact
        return abs(Fraction(float(fv)) - Fraction(float(cand))) <= Fraction(eb)
    diff = dt(fv) - dt(cand)  # reference: rounded in the value type
    return abs(diff) <= dt(eb)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("mode", [0, 1])
def test_granted_steps_honour_the_bound(dev, dt, mode):
    from pmsz_anonymous_amd import engine
    rng = np.random.default_rng(11 + mode + (dt == np.float64) * 2)
    # a known case: a bound near the smallest normal, so the step is subnormal
    # (its reciprocal is flushed on the device); nothing may be granted
    if dt == np.float32:
        f = np.array([6.983819e-36], np.float32)
        g = np.array([6.9726014e-36], np.float32)
        b = engine.test_step_budget(torch.from_numpy(f).to(dev), torch.from_numpy(g).to(dev),
                                    torch.zeros(1, dtype=torch.uint8, device=dev),
                                    2.1278843364108736e-38, 2, mode).cpu().numpy()
        assert int(b[0]) == 0, b
    granted = checked = cells = 0
    for q in (2, 3, 5, 8, 15):
        for _ in range(12):
            n = 128
            f, g, e = _group(rng, n, dt)
            d = rng.integers(0, q + 1, n).astype(np.uint8)
            tf = torch.from_numpy(f).to(dev)
            tg = torch.from_numpy(g).to(dev)
            td = torch.from_numpy(d).to(dev)
            b = engine.test_step_budget(tf, tg, td, e, q, mode).cpu().numpy()
            delta = dt(e / q)
            cells += n
            for i in range(n):
                di, bi = int(d[i]), int(b[i])
                assert bi <= q - 1, (bi, q)
                if bi <= di:
                    continue
                granted += 1
                gv = g[i]
                for step in range(di, bi):
                    cand = dt(gv - delta)
                    assert _accepted(f[i], cand, e, mode, dt), (
                        f"q={q} mode={mode} f={f[i]!r} g={g[i]!r} eb={e!r} d={di} B={bi} "
                        f"step {step}: {cand!r}")
                    gv = cand
                    checked += 1
    # the budgets are not vacuous: most cells inside the bound get some
    assert granted > cells // 4, (granted, cells)
    assert checked >= granted
