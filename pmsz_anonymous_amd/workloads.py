"""The benchmark workloads, shared by bench.py and the tests: synthetic 3-D
improved-Perlin fBm fields (fieldio.perlin) compressed by the host codecs, at the
sizes and parameters of the five configurations.

  C1  64^3 f32, seed 0, base frequency 4, 4 octaves, SZ-like, eb = 1e-3 x range
  C2  512^3 f32, seed 1, frequency 8, 5 octaves, SZ-like or ZFP, eb = 1e-3 x range
  C3  768^3 f64, seed 2, frequency 16, 6 octaves, SZ-like, eb = 1e-4 x range,
      corrected as 2 x 2 x 2 subdomains
  C4  64 f32 fields, edges log-uniform in [32, 256] per axis (rng 0), seeds
      100-163, eb log-uniform in [1e-4, 1e-2] x range, SZ-like / ZFP by seed parity
  C5  correction jobs: edges log-uniform in [64, 512] per axis, f32 80% / f64 20%,
      eb in {1e-2, 1e-3, 1e-4} x range, Perlin seed = the request counter
"""
from __future__ import annotations

import numpy as np

from . import fieldio


def c1_case():
    """(f, g0, eb) of C1."""
    return fieldio.perlin_case((64, 64, 64), 0, 4, 4, 1e-3, "sz3", np.float32)


def c2_case(compressor="sz3", edge=512):
    """(f, g0, eb) of C2 with the SZ-like codec ("sz3") or ZFP ("zfp")."""
    return fieldio.perlin_case((edge,) * 3, 1, 8.0, 5, 1e-3, compressor, np.float32)


def c3_case(edge=768):
    """(f, g0, eb) of C3 (the field; the 2 x 2 x 2 split is the caller's sub)."""
    return fieldio.perlin_case((edge,) * 3, 2, 16.0, 6, 1e-4, "sz3", np.float64)


def c4_fields(n=64, seed0=100, rng_seed=0):
    """C4: n tuples (f, g0, eb, compressor); base frequency 4, 4 octaves; ZFP only
    where libzfp loads (the SZ-like codec otherwise)."""
    rng = np.random.default_rng(rng_seed)
    out = []
    for k in range(n):
        edges = np.exp(rng.uniform(np.log(32), np.log(256), size=3)).round().astype(int)
        eb_rel = float(10 ** rng.uniform(-4, -2))
        seed = seed0 + k
        comp = "zfp" if seed % 2 and fieldio.zfp_available() else "sz3"
        shape = (int(edges[2]), int(edges[1]), int(edges[0]))
        out.append(fieldio.perlin_case(shape, seed, 4, 4, eb_rel, comp, np.float32) + (comp,))
    return out


def c5_jobs(n, rng_seed=0, first_seed=0):
    """C5: n jobs (f, g0, eb) for request counters first_seed.. (base frequency 8,
    5 octaves, SZ-like codec)."""
    rng = np.random.default_rng(rng_seed)
    out = []
    for i in range(n):
        edges = np.exp(rng.uniform(np.log(64), np.log(512), size=3)).round().astype(int)
        dt = np.float64 if rng.uniform() < 0.2 else np.float32
        eb_rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
        out.append(fieldio.perlin_case((int(edges[2]), int(edges[1]), int(edges[0])),
                                       first_seed + i, 8, 5, eb_rel, "sz3", dt))
    return out


def poisson_arrivals(n, rate, rng_seed=0):
    """Arrival times (seconds from 0) of n requests at mean rate `rate` per second."""
    rng = np.random.default_rng(rng_seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
