"""Target tags of the full-volume detections go on from one correction to the
next in a context, and the target byte map is cleared whole only when the tags
start over or its memory held something else.  Many corrections through one
context, each equal to the oracle's correction bit for bit:
* every run enqueues at least 9 detections, so 44 runs take the tags past 255 at
  least once, also in the middle of a run;
* fields grow and shrink, so stale marks lie beyond a run's cells;
* codec, edit-encoder and naive-mode calls in between use the map's memory as
  scratch."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STAT_KEYS = ("edit_sweeps", "check_sweeps", "ite", "total_targets", "clamped_voxels",
             "status", "final_cnt")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    return torch.device("cuda:0")


def _case(k):
    from pmsz_anonymous_amd import fieldio
    edges = (12, 20, 33, 48, 24, 40, 16)
    e = edges[k % len(edges)]
    shape = (e, e + k % 5, e + 3)
    dtype = np.float64 if k % 4 == 3 else np.float32
    eb_rel = (1e-3, 3e-3, 1e-2)[k % 3]
    return fieldio.perlin_case(shape, 700 + k, 4 + k % 3, 3, eb_rel, "sz3", dtype)


def _check(f, g0, eb, g, st, what):
    go, dco, sto, _, rc = O.correct(f, g0, eb)
    assert rc == 0, what
    assert g.tobytes() == go.tobytes(), f"{what}: field differs from the oracle"
    for key in STAT_KEYS:
        assert st[key] == sto[key], (what, key, st[key], sto[key])


@pytest.mark.parametrize("use_ctx", [False, True])
def test_tags_carry_over_runs(dev, use_ctx):
    from pmsz_anonymous_amd import engine
    ctx = engine.Context() if use_ctx else None
    try:
        for k in range(44):
            f, g0, eb = _case(k)
            tf = torch.from_numpy(f).to(dev)
            tg = torch.from_numpy(g0).to(dev)
            print(f"run {k}: shape {f.shape} {f.dtype} eb {eb:.3g}", flush=True)
            if ctx is None:
                _, st = engine.correct(tf, tg, eb)
            else:
                _, st = ctx.correct(tf, tg, eb)
            torch.cuda.synchronize()
            _check(f, g0, eb, tg.cpu().numpy(), st, f"run {k}")
            if ctx is None and k % 9 == 4:
                # the codec and the edit encoder use the default context's map as scratch
                engine.codec_dq(tf, eb)
                engine.encode_edits(tf, torch.from_numpy(g0).to(dev), tg, eb)
            if ctx is None and k % 13 == 6:
                tn = torch.from_numpy(g0).to(dev)
                engine.correct(tf, tn, eb, mode=engine.NAIVE)
    finally:
        if ctx is not None:
            ctx.close()
