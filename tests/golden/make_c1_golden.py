"""Writes tests/golden/c1_64cube.json: the C1 workload (64^3 f32 Perlin, seed 0,
base frequency 4, 4 octaves, SZ-like codec, eb = 1e-3 x range) and the oracle's
correction of it, as SHA-256 digests of the little-endian arrays plus the run's
counts, in both clamp modes.  Run from the repository root:

    python tests/golden/make_c1_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def record():
    import pyoracle as O
    from pmsz_anonymous_amd import workloads
    f, g0, eb = workloads.c1_case()
    asc, desc = O.labels(f)
    out = {"workload": "64^3 f32 Perlin (seed 0, freq 4, 4 octaves), SZ-like codec, "
                       "abs eb = 1e-3 x range",
           "shape": list(f.shape), "abs_eb": float(eb),
           "f": digest(f), "g0": digest(g0), "codes": digest(O.codes(f)),
           "labels_asc": digest(asc.astype(np.int64)),
           "labels_desc": digest(desc.astype(np.int64)), "runs": {}}
    for name, clamp in (("bound_safe", O.CLAMP_BOUND_SAFE), ("reference", O.CLAMP_REFERENCE)):
        g, dc, st, _, rc = O.correct(f, g0, eb, clamp_mode=clamp)
        assert rc == 0
        out["runs"][name] = {
            "g": digest(g), "dc": digest(dc), "edit_sweeps": st["edit_sweeps"],
            "check_sweeps": st["check_sweeps"], "ite": st["ite"],
            "total_targets": st["total_targets"], "clamped_voxels": st["clamped_voxels"],
            "edited": int((g < g0).sum())}
    return out


if __name__ == "__main__":
    rec = record()
    with open(os.path.join(HERE, "c1_64cube.json"), "w") as fh:
        json.dump(rec, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(json.dumps(rec["runs"]))
