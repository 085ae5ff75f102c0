"""Tuning and the autotuner: the schedule and tile parameters change the speed of
a run, never its result.  Every tuning gives the default tuning's corrected field
and statistics bit for bit; the autotuner leaves its inputs untouched and keeps a
valid tuning."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TUNINGS = [
    {"zchunk": 16},
    {"zchunk": 32, "dense_div": 4},
    {"dense_div": 64, "tail_targets": 1024, "tail_blocks": 16},
    {"tail_targets": 0},          # no tail: every sweep as separate launches
    {"tail_targets": 1 << 30, "tail_blocks": 128},  # the tail from the first sparse sweep
    {"dense_div": 1 << 12},       # full-volume detections only while targets are many
    {"solo_targets": 0},          # the tail never shrinks to one workgroup
    {"solo_targets": 1 << 20, "tail_blocks": 4},  # ... or from its first sweep
    {"stream_div": 0},            # targets of full-volume detections always listed
    {"stream_div": 1 << 24},      # ... streamed from the byte map whenever any violation
    {"stream_div": 64, "dense_div": 8},
    {"graphs": 1},                # multi-domain runs replay captured chains
    {"graphs": 1, "zchunk": 32, "tail_targets": 4096},
]
KEYS = ("edit_sweeps", "check_sweeps", "total_targets", "touched_voxels", "clamped_voxels",
        "final_cnt", "status", "trace_n")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from pmsz_anonymous_amd import engine
    engine.lib()
    yield torch.device("cuda:0")
    engine.set_tuning({})


def _run(f, g0, eb, **kw):
    from pmsz_anonymous_amd import engine
    g = g0.clone()
    _, st = engine.correct(f, g, eb, **kw)
    return g, st


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tunings_do_not_change_results(dev, dtype):
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((41, 70, 97), 11, 6, 4, 1e-3, "sz3", dtype)
    tf, tg0 = torch.from_numpy(f).to(dev), torch.from_numpy(g0).to(dev)
    engine.set_tuning({})
    ref, st0 = _run(tf, tg0, eb)
    for t in TUNINGS:
        engine.set_tuning(t)
        g, st = _run(tf, tg0, eb)
        assert torch.equal(g, ref), t
        for k in KEYS:
            assert st[k] == st0[k], (t, k)
        assert st["trace_targets"][:st["trace_n"]] == st0["trace_targets"][:st0["trace_n"]], t
    engine.set_tuning({})


@pytest.mark.parametrize("mode", [0, 1])
def test_tunings_subdomains_and_batch(dev, mode):
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((34, 52, 60), 12, 5, 4, 1e-3, "sz3", np.float32)
    tf, tg0 = torch.from_numpy(f).to(dev), torch.from_numpy(g0).to(dev)
    cases = [fieldio.perlin_case(s, 20 + k, 5, 3, e, "sz3", np.float32)
             for k, (s, e) in enumerate([((20, 33, 47), 1e-2), ((64, 24, 40), 1e-3),
                                         ((30, 30, 30), 1e-4)])]
    fs = [torch.from_numpy(c[0]).to(dev) for c in cases]
    g0s = [torch.from_numpy(c[1]).to(dev) for c in cases]
    ebs = [c[2] for c in cases]
    engine.set_tuning({})
    ref, st0 = _run(tf, tg0, eb, sub=(2, 2, 2), mode=mode)
    refb = [g.clone() for g in g0s]
    stb0 = engine.correct_batch(fs, refb, ebs, mode=mode)
    for t in TUNINGS[:4] + TUNINGS[-2:]:
        engine.set_tuning(t)
        g, st = _run(tf, tg0, eb, sub=(2, 2, 2), mode=mode)
        assert torch.equal(g, ref), t
        for k in ("edit_sweeps", "outer_iters", "total_targets", "status"):
            assert st[k] == st0[k], (t, k)
        gb = [g.clone() for g in g0s]
        stb = engine.correct_batch(fs, gb, ebs, mode=mode)
        for k in range(len(gb)):
            assert torch.equal(gb[k], refb[k]), (t, k)
            assert stb[k]["total_targets"] == stb0[k]["total_targets"], (t, k)
    engine.set_tuning({})


def test_autotune_single_and_batch(dev):
    from pmsz_anonymous_amd import engine, fieldio
    f, g0, eb = fieldio.perlin_case((48, 64, 80), 13, 6, 4, 1e-3, "sz3", np.float32)
    tf, tg0 = torch.from_numpy(f).to(dev), torch.from_numpy(g0).to(dev)
    keep = tg0.clone()
    engine.set_tuning({})
    ref, _ = _run(tf, tg0, eb)
    best, secs = engine.autotune([tf], [tg0], [eb], reps=1)
    assert torch.equal(tg0, keep), "the autotuner modified its input"
    assert secs > 0 and best["zchunk"] in (16, 32, 64) and best["dense_div"] >= 1
    assert engine.get_tuning() == best
    g, _ = _run(tf, tg0, eb)
    assert torch.equal(g, ref)
    # a batch in a context of its own
    cases = [fieldio.perlin_case((24 + 8 * k, 40, 36), 30 + k, 5, 3, 1e-3, "sz3", np.float32)
             for k in range(4)]
    fs = [torch.from_numpy(c[0]).to(dev) for c in cases]
    g0s = [torch.from_numpy(c[1]).to(dev) for c in cases]
    ebs = [c[2] for c in cases]
    ctx = engine.Context()
    try:
        refb = [g.clone() for g in g0s]
        ctx.correct_batch(fs, refb, ebs)
        best_b, _ = engine.autotune(fs, g0s, ebs, reps=1, ctx=ctx)
        assert engine.get_tuning(ctx) == best_b
        gb = [g.clone() for g in g0s]
        ctx.correct_batch(fs, gb, ebs)
        assert all(torch.equal(a, b) for a, b in zip(gb, refb))
    finally:
        ctx.close()
    engine.set_tuning({})


def test_graph_replays_match_eager(dev):
    # captured chains are replayed on later calls with the same layout: every
    # replay, with eager work of the program in between, equals the eager result
    from pmsz_anonymous_amd import engine, fieldio
    cases = [fieldio.perlin_case((20 + 6 * k, 36, 44), 40 + k, 5, 3, 10 ** (-2 - k % 3), "sz3",
                                 np.float32) for k in range(5)]
    fs = [torch.from_numpy(c[0]).to(dev) for c in cases]
    g0s = [torch.from_numpy(c[1]).to(dev) for c in cases]
    ebs = [c[2] for c in cases]
    ctx = engine.Context()
    try:
        engine.set_tuning({}, ctx)
        ref = [g.clone() for g in g0s]
        st_ref = ctx.correct_batch(fs, ref, ebs)
        engine.set_tuning({"graphs": 1}, ctx)
        x = torch.randn(1 << 20, device=dev)
        for rep in range(4):
            gs = [g.clone() for g in g0s]
            st = ctx.correct_batch(fs, gs, ebs)
            for k in range(len(gs)):
                assert torch.equal(gs[k], ref[k]), (rep, k)
                assert st[k]["edit_sweeps"] == st_ref[k]["edit_sweeps"], (rep, k)
            # eager work between replays: torch kernels and a plain correction
            x = torch.sort(x * 1.0001)[0]
            gp = g0s[0].clone()
            ctx.correct(fs[0], gp, ebs[0])
            torch.cuda.synchronize()
    finally:
        ctx.close()
