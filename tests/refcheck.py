"""Independent checkers used to pin the oracle: brute-force classification by
Python tuple comparison over an explicit neighbour list, pointer chasing, and
exact bound checks with fractions.  Written from the reference's definitions,
not from the oracle's code:

* stencil: the 14 (dx, dy, dz) triples of directions_host, pMSz.cu:186-193;
* order: value first, then linear index (islarger_shared / isless_shared,
  pMSz.cu:610-620) -- exactly Python's lexicographic tuple order on
  (value, index);
* a voxel is a maximum iff it is the largest of itself and its neighbours
  (pMSz.cu:709-717), a minimum likewise (pMSz.cu:719-725).
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

# directions_host, pMSz.cu:186-193, as (dx, dy, dz)
STENCIL = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
           (-1, 1, 0), (1, -1, 0), (-1, 0, 1), (1, 0, -1), (0, 1, 1), (0, -1, -1),
           (-1, 1, 1), (1, -1, -1)]

# reference direction index (0..13, 14 = the voxel) -> rank of the candidate by
# linear index (0..14, 7 = the voxel), the code layout of libpmsz.so
DIR2RANK = [8, 6, 10, 4, 12, 2, 9, 5, 11, 3, 14, 0, 13, 1, 7]


def dir_codes_to_rank(codes):
    t = np.array(DIR2RANK, np.uint8)
    lo = t[codes & 15]
    hi = t[codes >> 4]
    return (lo | (hi << 4)).astype(np.uint8)


def brute_classify(field):
    """Per voxel (asc_index, desc_index) by tuple max/min over the closed
    neighbourhood; returns two int64 arrays of global linear indices."""
    nz, ny, nx = field.shape
    flat = field.reshape(-1)
    asc = np.empty(flat.size, np.int64)
    desc = np.empty(flat.size, np.int64)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                cands = [(flat[i], i)]
                for dx, dy, dz in STENCIL:
                    a, b, c = x + dx, y + dy, z + dz
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        j = a + nx * (b + ny * c)
                        cands.append((flat[j], j))
                asc[i] = max(cands)[1]
                desc[i] = min(cands)[1]
    return asc.reshape(field.shape), desc.reshape(field.shape)


def chase(ptr):
    """Root of every pointer chain, by plain repeated following."""
    flat = ptr.reshape(-1)
    out = np.empty_like(flat)
    for i in range(flat.size):
        c = i
        while flat[c] != c:
            c = flat[c]
        out[i] = c
    return out.reshape(ptr.shape)


def exact_within(f, g, eb):
    """True iff |f - g| <= eb holds exactly for every voxel (rational arithmetic)."""
    e = Fraction(float(eb))
    for a, b in zip(np.asarray(f).reshape(-1).tolist(), np.asarray(g).reshape(-1).tolist()):
        if abs(Fraction(a) - Fraction(b)) > e:
            return False
    return True
