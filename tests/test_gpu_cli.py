"""The pmsz executable end to end, as the reference is run (pMSz.cu:2596-3188):
raw f64 input, SZ-style compression, correction over a 2x1x2 rank grid, the
corrected field and the edit files checked against the oracle, and the report
lines (iteration, storage overhead, CR, overall CR, edit ratio)."""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from editcheck import rank_blocks

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pmsz_anonymous_amd", "pmsz")


def test_cli_sz3_subdomains(tmp_path):
    from pmsz_anonymous_amd import fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    nz, ny, nx = 30, 36, 40
    f = fieldio.perlin((nz, ny, nx), 3, 5, 4, dtype=np.float64)
    lo, hi = fieldio.value_range(f)
    eb = 2e-3 * (hi - lo)
    src = tmp_path / "field.raw"
    f.tofile(src)
    out = tmp_path / "corrected.raw"
    cmd = [EXE, f"{src},{nx},{ny},{nz}", repr(eb), "sz3", "2", "1", "2", "--out", str(out),
           "--mode", "sync"]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    text = r.stdout

    g0, info = fieldio.szlike(f, eb)
    go, _, sto, _, rc = O.correct(f, g0, eb, sub=(2, 1, 2), mode=O.SYNC)
    assert rc == 0
    g = np.fromfile(out, np.float64).reshape(f.shape)
    assert g.tobytes() == go.tobytes()
    assert f"iteration is: {sto['ite']}" in text

    # the edit files in the working directory, prefix ""
    sym, vals = fieldio.read_edits(str(tmp_path) + "/", np.float64, 5,
                                   [f[sl].size for sl in rank_blocks(f.shape, (2, 1, 2))],
                                   f.size)
    want, edited = O.edit_symbols(f, g0, go, eb)
    assert sym.tobytes() == np.concatenate(
        [want[sl].ravel() for sl in rank_blocks(f.shape, (2, 1, 2))]).tobytes()
    _, ovals, _, _, _ = O.encode_edits(f, g0, go, eb, sub=(2, 1, 2))
    assert vals.tobytes() == ovals.tobytes()

    nd = os.path.getsize(tmp_path / "_deltas.bin")
    ne = os.path.getsize(tmp_path / "_edits.bin")
    assert re.search(rf"total storage overhead is: {nd + ne}\n", text)
    m = re.search(r"compression time: \S+ s \(sz3, (\d+) bytes\)", text)
    cmp_bytes = int(m.group(1))
    ocr = float(re.search(r"overall CR is: (\S+)", text).group(1))
    assert ocr == pytest.approx(f.size * 8 / (cmp_bytes + nd + ne), rel=1e-12)
    assert f"total_edited:{edited}" in text


def test_cli_naive_mode(tmp_path):
    from pmsz_anonymous_amd import fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    nz, ny, nx = 20, 24, 28
    f = fieldio.perlin((nz, ny, nx), 4, 5, 4, dtype=np.float64)
    lo, hi = fieldio.value_range(f)
    eb = 5e-3 * (hi - lo)
    src = tmp_path / "field.raw"
    f.tofile(src)
    out = tmp_path / "corrected.raw"
    cmd = [EXE, f"{src},{nx},{ny},{nz}", repr(eb), "sz3", "1", "1", "1", "--out", str(out),
           "--mode", "naive"]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g0, _ = fieldio.szlike(f, eb)
    go, _, sto, _, rc = O.correct(f, g0, eb, mode=O.NAIVE)
    assert rc == 0
    assert np.fromfile(out, np.float64).tobytes() == go.tobytes()
    assert f"iteration is: {sto['ite']}" in r.stdout


def test_cli_gpu_codec(tmp_path):
    from pmsz_anonymous_amd import fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    nz, ny, nx = 22, 26, 30
    f = fieldio.perlin((nz, ny, nx), 6, 5, 4, dtype=np.float64)
    lo, hi = fieldio.value_range(f)
    eb = 2e-3 * (hi - lo)
    src = tmp_path / "field.raw"
    f.tofile(src)
    out = tmp_path / "corrected.raw"
    cmd = [EXE, f"{src},{nx},{ny},{nz}", repr(eb), "dq", "2", "1", "1", "--out", str(out)]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g0, hist = O.codec_dq(f, eb, (2, 1, 1))
    go, _, sto, _, rc = O.correct(f, g0, eb, sub=(2, 1, 1))
    assert rc == 0
    assert np.fromfile(out, np.float64).tobytes() == go.tobytes()
    m = re.search(r"compression time: \S+ s \(dq, (\d+) bytes\)", r.stdout)
    # the size is the real stream's: the oracle's stream of the same field
    _, ostream = O.dq_compress(f, eb, (2, 1, 1))
    assert m and int(m.group(1)) == ostream.size
    assert "decoded back exactly" in r.stdout
    assert f"compressed size: {ostream.size} bytes (stream)" in r.stdout
    cr = float(re.search(r"original CR is: (\S+)", r.stdout).group(1))
    assert abs(cr - f.nbytes / ostream.size) < 1e-9 * cr


def test_cli_lazy_after_parent_cooperative_tail(tmp_path):
    # the parent runs a lazy correction (its tail is a cooperative launch) and keeps
    # its contexts; a child process's lazy subdomain run must still complete
    import torch
    from pmsz_anonymous_amd import engine, fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    f0, g00, eb0 = fieldio.perlin_case((40, 40, 40), 2, 5, 4, 3e-3, "sz3", np.float32)
    dev = torch.device("cuda:0")
    _, st = engine.correct(torch.from_numpy(f0).to(dev), torch.from_numpy(g00).to(dev), eb0)
    assert st["status"] == 0
    nz, ny, nx = 24, 28, 32
    f = fieldio.perlin((nz, ny, nx), 8, 5, 4, dtype=np.float64)
    lo, hi = fieldio.value_range(f)
    eb = 3e-3 * (hi - lo)
    src = tmp_path / "field.raw"
    f.tofile(src)
    out = tmp_path / "corrected.raw"
    cmd = [EXE, f"{src},{nx},{ny},{nz}", repr(eb), "sz3", "2", "2", "1", "--out", str(out)]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stderr, flush=True)  # a refused cooperative launch is reported here
    assert r.returncode == 0, r.stdout + r.stderr
    g0, _ = fieldio.szlike(f, eb)
    go, _, _, _, rc = O.correct(f, g0, eb, sub=(2, 2, 1))
    assert rc == 0
    assert np.fromfile(out, np.float64).tobytes() == go.tobytes()


def test_cli_readme_argument_form(tmp_path):
    # README.md:68-74: <path> <W> <H> <D> <abs_eb> <compressor> <PX> <PY> <PZ>
    from pmsz_anonymous_amd import fieldio
    if fieldio.zstd_version() is None:
        pytest.skip("libzstd absent")
    nz, ny, nx = 18, 20, 22
    f = fieldio.perlin((nz, ny, nx), 9, 5, 4, dtype=np.float64)
    lo, hi = fieldio.value_range(f)
    eb = 4e-3 * (hi - lo)
    src = tmp_path / "data.bin"
    f.tofile(src)
    out = tmp_path / "corrected.raw"
    cmd = [EXE, str(src), str(nx), str(ny), str(nz), repr(eb), "sz3", "1", "2", "1",
           "--out", str(out)]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g0, _ = fieldio.szlike(f, eb)
    go, _, sto, _, rc = O.correct(f, g0, eb, sub=(1, 2, 1))
    assert rc == 0
    assert np.fromfile(out, np.float64).tobytes() == go.tobytes()
    assert f"iteration is: {sto['ite']}" in r.stdout
