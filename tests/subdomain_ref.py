"""An independent, slow restatement of pMSz's subdomain correction, written from
the reference's text (not from the oracle's code), to pin the oracle's ghost
path: the padded blocks and their initial ghost values, the detection and edit
rules, the two-layer delta-counter exchange, and the lazy and per-iteration
loops.  Small fields only (pure Python loops).

Followed line by line:
* block split and padded blocks: pMSz.cu:2690-2711, read_subblock 274-313;
* initial ghost values: exchange_ghost_layers, pMSz.cu:320-608 -- every halo
  face of a direction in the loop (k < 12, pMSz.cu:332) is filled with
  INVALID_VAL, then the neighbour's core cells arrive for the faces and the
  edges that have a branch (pMSz.cu:394-488); the "full" policy of this build
  sends every stencil direction instead;
* classification: iscriticle, pMSz.cu:676-728 (value, then padded index);
  original pointers pMSz.cu:739-783; detection pMSz.cu:786-895 under snapshot
  semantics (each target edited once per sweep, on its start-of-sweep state);
* edit: applyDeltaBuffer_local, pMSz.cu:623-645, incl. send_flag (target in the
  outer two layers of the padded block);
* delta-counter exchange: get_send_count, pack_send_buffer_two_layers_kernel,
  unpack_recv_buffer_two_layers_kernel and their host loop,
  pMSz.cu:909-1272 (12 directions; the full policy adds the two stencil
  corners), the merge `g -= delta * (recv - cur)`, clamp when recv reaches q;
* loops: lazy c_loops pMSz.cu:1943-2070, per-iteration sync_pMSz.cu:2102-2200;
  "iteration is:" is rank 0's ite (pMSz.cu:2073).
"""
from __future__ import annotations

import numpy as np

import pyoracle as O

INVALID = 1e30
# directions_host / directions_host1, pMSz.cu:186-205, as (dx, dy, dz)
DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
        (-1, 1, 0), (1, -1, 0), (-1, 0, 1), (1, 0, -1), (0, 1, 1), (0, -1, -1),
        (-1, 1, 1), (1, -1, -1)]
# the branches of exchange_ghost_layers' send / receive lists (pMSz.cu:394-488)
REF_BRANCHES = {(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                (1, -1, 0), (-1, 1, 0), (0, -1, 1), (0, 1, -1), (1, 0, -1), (-1, 0, 1)}


def split(n, parts, coord):
    """pMSz.cu:2690-2706: the first n % parts blocks get one more cell."""
    base, rem = divmod(n, parts)
    local = base + (1 if coord < rem else 0)
    off = coord * base + min(coord, rem)
    return local, off


class Block:
    def __init__(self, c, l, o, dtype):
        self.c, self.l, self.o = c, l, o          # (cx, cy, cz), (lx, ly, lz), (ox, oy, oz)
        self.P = tuple(v + 2 for v in l)          # padded (PX, PY, PZ)
        shp = (self.P[2], self.P[1], self.P[0])
        self.f = np.full(shp, INVALID, dtype)
        self.g = np.full(shp, INVALID, dtype)
        self.dc = np.zeros(shp, np.uint8)
        self.send_flag = 0
        self.ite = 0
        self.cnt = [1, 1, 1]                      # host_count_f_* start at 1 (pMSz.cu:1924)
        self.orig = {}

    def inside(self, x, y, z):
        return 0 <= x < self.P[0] and 0 <= y < self.P[1] and 0 <= z < self.P[2]

    def idx(self, x, y, z):
        return x + (y + z * self.P[1]) * self.P[0]


def make_blocks(f, g, sub, policy):
    nz, ny, nx = f.shape
    blocks = {}
    for cz in range(sub[2]):
        for cy in range(sub[1]):
            for cx in range(sub[0]):
                lx, ox = split(nx, sub[0], cx)
                ly, oy = split(ny, sub[1], cy)
                lz, oz = split(nz, sub[2], cz)
                b = Block((cx, cy, cz), (lx, ly, lz), (ox, oy, oz), f.dtype)
                b.f[1:lz + 1, 1:ly + 1, 1:lx + 1] = f[oz:oz + lz, oy:oy + ly, ox:ox + lx]
                b.g[1:lz + 1, 1:ly + 1, 1:lx + 1] = g[oz:oz + lz, oy:oy + ly, ox:ox + lx]
                blocks[(cx, cy, cz)] = b
    for b in blocks.values():
        for arr in ("f", "g"):
            initial_ghosts(blocks, b, arr, sub, policy)
    for b in blocks.values():
        label_original(b)
    return blocks


def initial_ghosts(blocks, b, arr, sub, policy):
    """exchange_ghost_layers: the halo cells a branch receives into hold the
    neighbour's core value (read through global coordinates); the rest of the
    halo stays INVALID_VAL."""
    data = getattr(b, arr)
    lx, ly, lz = b.l
    dirs = DIRS[:12] if policy == O.GHOST_REFERENCE else DIRS
    for d in dirs:
        if policy == O.GHOST_REFERENCE and d not in REF_BRANCHES:
            continue
        nb = tuple(b.c[a] + d[a] for a in range(3))
        if nb not in blocks:
            continue
        nbk = blocks[nb]
        # receive region of direction d (pMSz.cu:504-597): the halo coordinate along
        # each nonzero component, the core range along the others
        rng = []
        for a, L in enumerate((lx, ly, lz)):
            if d[a] == -1:
                rng.append([0])
            elif d[a] == 1:
                rng.append([L + 1])
            else:
                rng.append(range(1, L + 1))
        src = getattr(nbk, arr)
        for z in rng[2]:
            for y in rng[1]:
                for x in rng[0]:
                    gx, gy, gz = b.o[0] + x - 1, b.o[1] + y - 1, b.o[2] + z - 1
                    sx, sy, sz = gx - nbk.o[0] + 1, gy - nbk.o[1] + 1, gz - nbk.o[2] + 1
                    data[z, y, x] = src[sz, sy, sx]


def classify(b, h, x, y, z):
    """(is_max, is_min, largest, smallest) over the padded neighbourhood
    (pMSz.cu:676-728): tuple order on (value, padded index)."""
    i = b.idx(x, y, z)
    v = h[z, y, x]
    best = (v, i)
    worst = (v, i)
    is_max = is_min = True
    for dx, dy, dz in DIRS:
        nx_, ny_, nz_ = x + dx, y + dy, z + dz
        if not b.inside(nx_, ny_, nz_):
            continue
        nv = h[nz_, ny_, nx_]
        if nv == INVALID:
            continue
        key = (nv, b.idx(nx_, ny_, nz_))
        if key > (v, i):
            is_max = False
        if key < (v, i):
            is_min = False
        best = max(best, key)
        worst = min(worst, key)
    return is_max, is_min, best[1], worst[1]


def core_cells(b):
    lx, ly, lz = b.l
    for z in range(1, lz + 1):
        for y in range(1, ly + 1):
            for x in range(1, lx + 1):
                yield x, y, z


def label_original(b):
    """or_types (pMSz.cu:739-783): (is_max, is_min, asc, desc); a voxel that is
    both a maximum and a minimum counts as a maximum only."""
    for x, y, z in core_cells(b):
        if b.f[z, y, x] == INVALID:
            continue
        mx, mn, li, si = classify(b, b.f, x, y, z)
        i = b.idx(x, y, z)
        if mx:
            b.orig[i] = (True, False, i, si)
        elif mn:
            b.orig[i] = (False, True, li, i)
        else:
            b.orig[i] = (False, False, li, si)


def coords(b, i):
    PX, PY = b.P[0], b.P[1]
    return i % PX, (i // PX) % PY, i // (PX * PY)


def edit(b, t, eb, q, clamp_mode):
    """applyDeltaBuffer_local (pMSz.cu:623-645) in the value type."""
    T = b.g.dtype.type
    x, y, z = coords(b, t)
    delta, eps = T(eb / q), T(eb)
    fv, gv = b.f[z, y, x], b.g[z, y, x]
    cand = T(gv - delta)
    if clamp_mode == O.CLAMP_BOUND_SAFE:
        ok = O.within(fv, cand, eb)
    else:
        ok = abs(T(fv - cand)) <= eps
    if ok and b.dc[z, y, x] < q - 1:
        b.g[z, y, x] = cand
        b.dc[z, y, x] += 1
    else:
        b.dc[z, y, x] = q
        b.g[z, y, x] = (T(fv - eps) if clamp_mode == O.CLAMP_REFERENCE
                        else O.clamp(fv, eb, clamp_mode, b.g.dtype))
    PX, PY, PZ = b.P
    if x <= 1 or x >= PX - 2 or y <= 1 or y >= PY - 2 or z <= 1 or z >= PZ - 2:
        b.send_flag = 1


def sweep(b, eb, q, clamp_mode, check):
    """One iscriticle launch (data_type 0) on block b: counters, and unless check,
    one edit per distinct target on the start-of-sweep state."""
    cmax = cmin = cdir = 0
    targets = []
    for x, y, z in core_cells(b):
        if b.g[z, y, x] == INVALID:
            continue
        i = b.idx(x, y, z)
        omax, omin, oasc, odesc = b.orig[i]
        is_max, is_min, li, si = classify(b, b.g, x, y, z)
        fm = is_max != omax
        fmin = is_min != omin
        if fm:
            cmax += 1
            targets.append(li if omax else i)
        if fmin:
            cmin += 1
            targets.append(i if omin else odesc)
        if fm or fmin:
            cdir += 1
            continue
        if li != oasc:
            cdir += 1
            targets.append(li)
        if si != odesc:
            cdir += 1
            targets.append(odesc)
    uniq = sorted(set(targets))
    if not check:
        for t in uniq:
            edit(b, t, eb, q, clamp_mode)
    b.cnt = [cmax, cmin, cdir]
    return len(uniq) if not check else 0


def pack_cell(b, d, idx, recv):
    """Cell of entry idx of the two-layer buffer toward direction d: the sender's
    layout (pMSz.cu:949-985) or the receiver's (pMSz.cu:1027-1072)."""
    dx, dy, dz = d
    lx, ly, lz = b.l
    i, first = idx // 2, idx % 2 == 0

    def axis(dc, L):
        if not recv:
            return (1 - (0 if first else 1)) if dc == -1 else (L + (0 if first else 1))
        return (0 + (0 if first else 1)) if dc == -1 else (L + (1 if first else 0))

    if dx != 0 and dy == 0 and dz == 0:
        return axis(dx, lx), 1 + i % ly, 1 + i // ly
    if dy != 0 and dx == 0 and dz == 0:
        return 1 + i % lx, axis(dy, ly), 1 + i // lx
    if dz != 0 and dx == 0 and dy == 0:
        return 1 + i % lx, 1 + i // lx, axis(dz, lz)
    if dz == 0:
        return axis(dx, lx), axis(dy, ly), 1 + i
    if dy == 0:
        return axis(dx, lx), 1 + i, axis(dz, lz)
    if dx == 0:
        return 1 + i, axis(dy, ly), axis(dz, lz)
    return axis(dx, lx), axis(dy, ly), axis(dz, lz)   # corner: one cell per layer


def send_count(d, l):
    """get_send_count (pMSz.cu:909-925); a corner (full policy) sends 2 cells."""
    s = sum(abs(c) for c in d)
    lx, ly, lz = l
    if s == 1:
        return 2 * (ly * lz if d[0] else (lx * lz if d[1] else lx * ly))
    if s == 2:
        return 2 * (lx if d[0] == 0 else (ly if d[1] == 0 else lz))
    return 2


def delta_sync(blocks, shape, sub, eb, q, policy, clamp_mode):
    nz, ny, nx = shape
    dirs = DIRS[:12] if policy == O.GHOST_REFERENCE else DIRS
    bufs = {}
    for c, b in blocks.items():          # every rank packs before any unpack
        for d in dirs:
            nb = tuple(c[a] + d[a] for a in range(3))
            if nb not in blocks or not b.send_flag:
                continue
            n = send_count(d, b.l)
            buf = np.zeros(n, np.uint8)
            for idx in range(n):
                x, y, z = pack_cell(b, d, idx, recv=False)
                gx, gy, gz = x - 1 + b.o[0], y - 1 + b.o[1], z - 1 + b.o[2]
                if not (0 <= gx < nx and 0 <= gy < ny and 0 <= gz < nz):
                    continue
                if b.g[z, y, x] == INVALID:
                    continue
                if b.dc[z, y, x] != 0:
                    buf[idx] = b.dc[z, y, x]
            bufs[(c, d)] = buf
    for c, b in blocks.items():
        T = b.g.dtype.type
        delta = T(eb / q)
        for d in dirs:
            nb = tuple(c[a] + d[a] for a in range(3))
            if nb not in blocks:
                continue
            buf = bufs.get((nb, tuple(-v for v in d)))
            if buf is None:
                continue
            for idx in range(len(buf)):
                r = int(buf[idx])
                if r == 0:
                    continue
                x, y, z = pack_cell(b, d, idx, recv=True)
                cur = int(b.dc[z, y, x])
                if b.g[z, y, x] != INVALID and cur < r:
                    b.dc[z, y, x] = r
                    if r < q:
                        b.g[z, y, x] = T(b.g[z, y, x] - T(delta * T(r - cur)))
                    else:
                        eps = T(eb)
                        fv = b.f[z, y, x]
                        b.g[z, y, x] = (T(fv - eps) if clamp_mode == O.CLAMP_REFERENCE
                                        else O.clamp(fv, eb, clamp_mode, b.g.dtype))
    for b in blocks.values():
        b.send_flag = 0


def correct(f, g0, eb, sub, mode=O.LAZY, policy=O.GHOST_FULL, q=5,
            clamp_mode=O.CLAMP_BOUND_SAFE, max_outer=1000):
    """Returns (g, dc, stats) with edit_sweeps, check_sweeps, outer_iters, ite,
    total_targets."""
    blocks = make_blocks(f, g0, sub, policy)
    order = sorted(blocks, key=lambda c: (c[2], c[1], c[0]))
    st = dict(edit_sweeps=0, check_sweeps=0, outer_iters=0, total_targets=0)
    gl = [1, 1, 0]
    while any(gl):
        assert st["outer_iters"] < max_outer, "no convergence"
        if mode == O.LAZY:
            for c in order:
                b = blocks[c]
                b.ite += 1
                while any(b.cnt):
                    b.ite += 1
                    st["total_targets"] += sweep(b, eb, q, clamp_mode, False)
                    st["edit_sweeps"] += 1
        else:
            for c in order:
                b = blocks[c]
                b.ite += 2
                st["total_targets"] += sweep(b, eb, q, clamp_mode, False)
                st["edit_sweeps"] += 1
        if len(blocks) > 1:
            delta_sync(blocks, f.shape, sub, eb, q, policy, clamp_mode)
        for b in blocks.values():
            b.send_flag = 0
        gl = [0, 0, 0]
        for c in order:
            b = blocks[c]
            sweep(b, eb, q, clamp_mode, True)
            st["check_sweeps"] += 1
            gl = [max(gl[k], b.cnt[k]) for k in range(3)]
        st["outer_iters"] += 1
    st["ite"] = blocks[(0, 0, 0)].ite
    g = np.empty_like(g0)
    dc = np.zeros(f.shape, np.uint8)
    for b in blocks.values():
        (lx, ly, lz), (ox, oy, oz) = b.l, b.o
        g[oz:oz + lz, oy:oy + ly, ox:ox + lx] = b.g[1:lz + 1, 1:ly + 1, 1:lx + 1]
        dc[oz:oz + lz, oy:oy + ly, ox:ox + lx] = b.dc[1:lz + 1, 1:ly + 1, 1:lx + 1]
    return g, dc, st
