"""The serving loop (continuous batching over lanes): every job's result is the
same corrected alone and in the mix, bit for bit, for mixed value types, shapes
and bounds arriving while other jobs run; admission honours the HBM budget."""
import threading
import time

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


def _jobs(n, seed):
    from pmsz_anonymous_amd import fieldio
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        shape = tuple(int(v) for v in np.exp(rng.uniform(np.log(12), np.log(48), 3)).round())
        dtype = np.float64 if rng.uniform() < 0.2 else np.float32
        eb_rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
        out.append(fieldio.perlin_case(shape, 1000 + seed * 100 + k, 5, 3, eb_rel, "sz3", dtype))
    return out


def test_served_equals_alone(dev):
    from pmsz_anonymous_amd import engine, serve
    cases = _jobs(24, 1)
    alone = []
    for f, g0, eb in cases:
        tg = torch.from_numpy(g0).to(dev)
        _, st = engine.correct(torch.from_numpy(f).to(dev), tg, eb)
        alone.append((tg.cpu().numpy(), st))
    fs = [torch.from_numpy(f).to(dev) for f, _, _ in cases]
    gs = [torch.from_numpy(g0).to(dev) for _, g0, _ in cases]
    torch.cuda.synchronize()
    rng = np.random.default_rng(2)
    with serve.Server(lanes=3, max_batch=5) as srv:
        jobs = []
        for k, (f, g, (_, _, eb)) in enumerate(zip(fs, gs, cases)):
            jobs.append(srv.submit(f, g, eb))
            if k >= 12:  # a burst of 12, then arrivals while others run
                time.sleep(float(rng.exponential(0.002)))
        for j in jobs:
            j.wait(60)
        assert srv.runs < len(jobs), "no job was batched with another"
    for k, j in enumerate(jobs):
        got = j.g.cpu().numpy()
        want, st = alone[k]
        assert got.tobytes() == want.tobytes(), f"job {k} differs in the mix"
        for key in ("edit_sweeps", "total_targets", "touched_voxels", "clamped_voxels",
                    "final_cnt", "status"):
            assert j.stats[key] == st[key], (k, key)


def test_two_lanes_overlap_results(dev):
    # the same field submitted many times from two threads: all results identical
    from pmsz_anonymous_amd import engine, serve
    f, g0, eb = _jobs(1, 3)[0]
    tf = torch.from_numpy(f).to(dev)
    ref = torch.from_numpy(g0).to(dev)
    engine.correct(tf, ref, eb)
    copies = [torch.from_numpy(g0).to(dev) for _ in range(16)]
    torch.cuda.synchronize()
    with serve.Server(lanes=2, max_batch=4) as srv:
        jobs = []

        def client(part):
            for g in part:
                jobs.append(srv.submit(tf, g, eb))

        ts = [threading.Thread(target=client, args=(copies[i::2],)) for i in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for j in jobs:
            j.wait(60)
    for g in copies:
        assert torch.equal(g, ref)


def test_budget_admission(dev):
    from pmsz_anonymous_amd import engine, serve
    cases = _jobs(6, 4)
    fs = [torch.from_numpy(f).to(dev) for f, _, _ in cases]
    gs = [torch.from_numpy(g0).to(dev) for _, g0, _ in cases]
    need = max(engine.workspace_bytes([(tuple(f.shape), f.dtype)]) for f in fs)
    with serve.Server(lanes=2, hbm_budget=need * 2, max_batch=8) as srv:
        with pytest.raises(ValueError):
            big = torch.zeros((300, 300, 300), dtype=torch.float32, device=dev)
            srv.submit(big, big.clone(), 1e-3)
        jobs = [srv.submit(f, g, eb) for f, g, (_, _, eb) in zip(fs, gs, cases)]
        for j in jobs:
            j.wait(60)
    for j, (f, g0, eb) in zip(jobs, cases):
        tg = torch.from_numpy(g0).to(dev)
        engine.correct(torch.from_numpy(f).to(dev), tg, eb)
        assert torch.equal(tg, j.g)


def test_large_jobs_run_alone(dev):
    # jobs at or above alone_voxels take a run of their own; smaller ones batch
    from pmsz_anonymous_amd import engine, serve
    cases = _jobs(12, 5)
    fs = [torch.from_numpy(f).to(dev) for f, _, _ in cases]
    gs = [torch.from_numpy(g0).to(dev) for _, g0, _ in cases]
    sizes = sorted(f.numel() for f in fs)
    cut = sizes[len(sizes) // 2]
    torch.cuda.synchronize()
    with serve.Server(lanes=2, max_batch=8, alone_voxels=cut) as srv:
        jobs = [srv.submit(f, g, eb) for f, g, (_, _, eb) in zip(fs, gs, cases)]
        for j in jobs:
            j.wait(60)
    for j, f in zip(jobs, fs):
        if f.numel() >= cut:
            assert j.batch_size == 1
    for j, (f, g0, eb) in zip(jobs, cases):
        tg = torch.from_numpy(g0).to(dev)
        engine.correct(torch.from_numpy(f).to(dev), tg, eb)
        assert torch.equal(tg, j.g)
