"""The C-ABI libraries load and export every function their headers declare
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pmsz_anonymous_amd")


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pmsz_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("lib,header", [("libpmsz.so", "pmsz.h"), ("libpmsz_io.so", "pmsz_io.h")])
def test_exports_match_header(lib, header):
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), f"{lib} not built"
    L = ctypes.CDLL(path)
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_engine_host_helpers_without_device():
    from pmsz_anonymous_amd import engine
    p = engine.make_params((4, 5, 6), "float32", 1e-3)
    assert (p.nx, p.ny, p.nz) == (6, 5, 4) and p.q == 5 and p.max_sweeps == 100000
    assert engine.version().startswith("pmsz")
    assert "not converged" in engine.lib().pmsz_strerror(engine.NOT_CONVERGED).decode()


def test_bad_params_rejected_before_device_use():
    from pmsz_anonymous_amd import engine
    L = engine.lib()
    for kw in ({"abs_eb": 0.0}, {"abs_eb": float("nan")}, {"q": 0}, {"max_sweeps": 0}):
        p = engine.make_params((2, 2, 2), "float32", kw.get("abs_eb", 1e-3),
                               q=kw.get("q", 5), max_sweeps=kw.get("max_sweeps", 10))
        st = engine.Stats()
        rc = L.pmsz_correct(ctypes.byref(p), ctypes.c_void_p(8), ctypes.c_void_p(8), None,
                            ctypes.byref(st))
        assert rc == engine.EINVAL
    p = engine.make_params((1, 1, 1), "float32", 1e-3)
    assert L.pmsz_correct(ctypes.byref(p), ctypes.c_void_p(8), ctypes.c_void_p(8), None,
                          None) == engine.EINVAL


def test_tuning_defaults_and_validation():
    from pmsz_anonymous_amd import engine
    L = engine.lib()
    t = engine.Tuning()
    L.pmsz_default_tuning(ctypes.byref(t))
    assert t.as_dict() == {"dense_div": 32, "tail_targets": 4096, "tail_blocks": 32,
                           "solo_targets": 64, "zchunk": 64, "graphs": 0,
                           "stream_div": 32}
    for bad in ({"zchunk": 48}, {"zchunk": 128}, {"dense_div": 0}, {"tail_blocks": 0},
                {"tail_blocks": 257}, {"tail_targets": -1}, {"graphs": 2}, {"stream_div": -1},
                {"solo_targets": -1}):
        with pytest.raises(engine.PmszError) as e:
            engine.set_tuning(bad)
        assert e.value.code == engine.EINVAL
    assert engine.get_tuning() == t.as_dict()  # a rejected tuning leaves the old one


def test_io_codecs_bound_and_determinism():
    import numpy as np
    from pmsz_anonymous_amd import fieldio
    f = fieldio.perlin((16, 17, 18), 0, 4, 4)
    assert f.dtype == np.float32 and -1.5 < f.min() < f.max() < 1.5
    assert np.array_equal(f, fieldio.perlin((16, 17, 18), 0, 4, 4))
    lo, hi = fieldio.value_range(f)
    eb = 1e-3 * (hi - lo)
    g0, info = fieldio.szlike(f, eb)
    assert fieldio.bound_violations(f, g0, eb)[1] == 0
    if fieldio.zfp_available():
        gz, zi = fieldio.zfp(f, eb)
        assert fieldio.bound_violations(f, gz, eb)[1] == 0 and zi["bytes"] > 0
