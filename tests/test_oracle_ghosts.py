"""Pins of the oracle's subdomain path (no GPU): the initial ghost values, the
two-layer delta-counter exchange and the lazy / per-iteration loops.

* Initial ghost values against a closed form read off exchange_ghost_layers
  (pMSz.cu:320-608): a halo cell in direction s of its block holds the global
  field's value iff the neighbouring block in direction s exists and s is one of
  the directions that function receives -- the 6 faces and the edges
  (-1,1,0), (1,-1,0), (-1,0,1), (1,0,-1) (its loop runs over directions 0..11,
  pMSz.cu:332, and (0,1,1) / (0,-1,-1) have no branch, pMSz.cu:394-488) -- or,
  under the full policy, any of the 14 stencil directions; every other halo cell
  holds INVALID_VAL (1e30).
* The delta-counter exchange against hand-worked states: the sender's outermost
  core layer goes to the receiver's ghost layer and its ghost layer to the
  receiver's core (pMSz.cu:949-985, 1027-1072); a larger counter r replaces cur
  and lowers g by one product delta * (r - cur) (pMSz.cu:1096-1099), r >= q
  clamps (pMSz.cu:1100), a sender whose send_flag is 0 sends nothing
  (pMSz.cu:1161-1179), sentinel cells are left alone (pMSz.cu:1094).
* The loops against hand-worked runs on a 1-D field of two blocks (ite,
  outer_iters, edit_sweeps, the corrected field), and against an independent
  restatement of the whole subdomain path (tests/subdomain_ref.py) on random
  fields with ties, both policies, both loops, uneven and thin splits.
"""
import numpy as np
import pytest

import pyoracle as O
import subdomain_ref as S

INV = 1e30
REF_DIRS = {(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (-1, 1, 0), (1, -1, 0), (-1, 0, 1), (1, 0, -1)}
FULL_DIRS = REF_DIRS | {(0, 1, 1), (0, -1, -1), (-1, 1, 1), (1, -1, -1)}


def _field(shape, seed, dtype, alphabet=None):
    rng = np.random.default_rng(seed)
    if alphabet:
        return rng.integers(0, alphabet, size=shape).astype(dtype)
    return rng.standard_normal(shape).astype(dtype)


# ---------------------------------------------------------------- initial ghosts
@pytest.mark.parametrize("policy", [O.GHOST_REFERENCE, O.GHOST_FULL])
@pytest.mark.parametrize("sub,shape", [((2, 2, 2), (5, 6, 7)), ((3, 2, 1), (4, 5, 7)),
                                       ((1, 3, 2), (6, 7, 3)), ((2, 2, 3), (7, 4, 5))])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_initial_ghosts_closed_form(policy, sub, shape, dtype):
    f = _field(shape, 1, dtype)
    g = _field(shape, 2, dtype)
    fb, gb = O.blocks(f, g, sub, ghost_policy=policy)
    dirs = REF_DIRS if policy == O.GHOST_REFERENCE else FULL_DIRS
    nz, ny, nx = shape
    for (pshape, (oz, oy, ox)), bf, bg in zip(O.block_shapes(shape, sub), fb, gb):
        PZ, PY, PX = pshape
        for z in range(PZ):
            for y in range(PY):
                for x in range(PX):
                    gx, gy, gz = ox + x - 1, oy + y - 1, oz + z - 1
                    s = tuple(-1 if c == 0 else (1 if c == L - 1 else 0)
                              for c, L in ((x, PX), (y, PY), (z, PZ)))
                    inside = 0 <= gx < nx and 0 <= gy < ny and 0 <= gz < nz
                    if s == (0, 0, 0):
                        want = (f[gz, gy, gx], g[gz, gy, gx])
                    elif inside and s in dirs:
                        want = (f[gz, gy, gx], g[gz, gy, gx])
                    else:
                        want = (dtype(INV), dtype(INV))
                    assert (bf[z, y, x], bg[z, y, x]) == want, (pshape, (x, y, z), s)


def test_initial_ghosts_hand_cells():
    # 4 x 4 x 4 field, 2 x 2 x 2 blocks of 2^3 (padded 4^3); block k = cx + 2 (cy + 2 cz)
    f = np.arange(64, dtype=np.float64).reshape(4, 4, 4)   # value = global index
    ref, _ = O.blocks(f, f, (2, 2, 2), ghost_policy=O.GHOST_REFERENCE)
    full, _ = O.blocks(f, f, (2, 2, 2), ghost_policy=O.GHOST_FULL)
    gi = lambda x, y, z: x + 4 * (y + 4 * z)  # noqa: E731
    # block (1,0,0), padded (0, 3, 1): direction (-1, 1, 0) toward block (0,1,0),
    # global (1, 2, 0): exchanged under both policies
    assert ref[1][1, 3, 0] == gi(1, 2, 0) and full[1][1, 3, 0] == gi(1, 2, 0)
    # block (0,0,0), padded (3, 0, 1): direction (1, -1, 0) -- no block there
    assert ref[0][1, 0, 3] == INV and full[0][1, 0, 3] == INV
    # block (0,0,0), padded (1, 3, 3): direction (0, 1, 1) toward block (0,1,1),
    # global (0, 2, 2): dropped by the reference, exchanged by the full policy
    assert ref[0][3, 3, 1] == INV and full[0][3, 3, 1] == gi(0, 2, 2)
    # block (0,1,1), padded (1, 0, 0): direction (0, -1, -1), global (0, 1, 1)
    assert ref[6][0, 0, 1] == INV and full[6][0, 0, 1] == gi(0, 1, 1)
    # block (1,0,0), padded (0, 3, 3): the stencil corner (-1, 1, 1), global (1, 2, 2)
    assert ref[1][3, 3, 0] == INV and full[1][3, 3, 0] == gi(1, 2, 2)
    # block (0,0,0), padded (3, 3, 3): (1, 1, 1) is not a stencil direction
    assert ref[0][3, 3, 3] == INV and full[0][3, 3, 3] == INV
    # block (0,0,0), padded (3, 3, 1): (1, 1, 0) is not a stencil direction
    assert ref[0][1, 3, 3] == INV and full[0][1, 3, 3] == INV
    # block (1,1,0), padded (0, 1, 3): direction (-1, 0, 1) toward block (0,1,1),
    # global (1, 2, 2): exchanged under both policies
    assert ref[3][3, 1, 0] == gi(1, 2, 2) and full[3][3, 1, 0] == gi(1, 2, 2)


@pytest.mark.parametrize("policy", [O.GHOST_REFERENCE, O.GHOST_FULL])
def test_restatement_blocks_equal_oracle(policy):
    f = _field((5, 6, 7), 3, np.float64)
    g = _field((5, 6, 7), 4, np.float64)
    fb, gb = O.blocks(f, g, (2, 3, 2), ghost_policy=policy)
    rb = S.make_blocks(f, g, (2, 3, 2), policy)
    order = sorted(rb, key=lambda c: (c[2], c[1], c[0]))
    for k, c in enumerate(order):
        assert rb[c].f.tobytes() == fb[k].tobytes()
        assert rb[c].g.tobytes() == gb[k].tobytes()


# ---------------------------------------------------------------- delta-counter sync
def _two_blocks_x(dtype):
    """(nz, ny, nx) = (2, 3, 4) split 2 x 1 x 1: blocks of lx = 2, padded (4, 5, 4).
    Block 0: global x 0, 1 at padded x 1, 2; its ghost x = 3 is global x 2.
    Block 1: global x 2, 3 at padded x 1, 2; its ghost x = 0 is global x 1."""
    shape = (2, 3, 4)
    f = np.full(shape, 5.0, dtype)
    g = np.full(shape, 4.75, dtype)
    g[0, 1, 1] = 0.3   # shared cell A: block 0 core (2,2,1), block 1 ghost (0,2,1)
    g[0, 2, 2] = 1.7   # shared cell B: block 0 ghost (3,3,1), block 1 core (1,3,1)
    g[0, 1, 2] = 2.3   # shared cell G: block 1 core (1,2,1), block 0 ghost (3,2,1)
    return shape, f, g


def _sync_both(shape, eb, sub, fb, gb, dcb, send, policy=O.GHOST_FULL,
               clamp_mode=O.CLAMP_REFERENCE, q=5):
    """The oracle's sync and the restatement's on the same padded blocks."""
    go, do = O.sync_blocks(shape, eb, sub, fb, gb, dcb, send, q=q, ghost_policy=policy,
                           clamp_mode=clamp_mode)
    f = np.zeros(shape, fb[0].dtype)
    rb = S.make_blocks(f, f, sub, policy)
    order = sorted(rb, key=lambda c: (c[2], c[1], c[0]))
    for k, c in enumerate(order):
        rb[c].f[...] = fb[k]
        rb[c].g[...] = gb[k]
        rb[c].dc[...] = dcb[k]
        rb[c].send_flag = send[k]
    S.delta_sync(rb, shape, sub, eb, q, policy, clamp_mode)
    for k, c in enumerate(order):
        assert rb[c].g.tobytes() == go[k].tobytes(), c
        assert rb[c].dc.tobytes() == do[k].tobytes(), c
    return go, do


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sync_kat_layers_product_max_merge_clamp_gating(dtype):
    T = np.dtype(dtype).type
    shape, f, g = _two_blocks_x(dtype)
    eb, q = 0.35, 5
    delta = T(eb / q)
    fb, gb = O.blocks(f, g, (2, 1, 1))
    dcb = [np.zeros_like(b, dtype=np.uint8) for b in fb]
    # A: block 0's outermost core cell (layer 0) with dc 3
    dcb[0][1, 2, 2] = 3
    # B: block 0's ghost cell (layer 1) with dc 4; the owner (block 1 core) has dc 1
    dcb[0][1, 3, 3] = 4
    dcb[1][1, 3, 1] = 1
    # C: block 0's ghost with dc 2 while the owner already has 3: no change
    dcb[0][2, 1, 3] = 2
    dcb[1][2, 1, 1] = 3
    # D: block 0's core cell at q = 5: the receiver's ghost is clamped
    dcb[0][2, 3, 2] = 5
    # E: block 1's core cell with dc 2, but block 1's send_flag is 0
    dcb[1][1, 1, 1] = 2
    # F: block 1's ghost holding the sentinel is never changed
    dcb[0][2, 2, 2] = 2
    gb[1][2, 2, 0] = T(INV)
    g_before = [b.copy() for b in gb]
    go, do = _sync_both(shape, eb, (2, 1, 1), fb, gb, dcb, [1, 0])
    want_g = [b.copy() for b in g_before]
    want_dc = [b.copy() for b in dcb]
    # A: g = 0.3 - fl(delta * 3), one product, not three subtractions
    a = T(T(0.3) - T(delta * T(3)))
    assert a != T(T(T(T(0.3) - delta) - delta) - delta)
    want_g[1][1, 2, 0] = a
    want_dc[1][1, 2, 0] = 3
    # B: cur 1 -> 4: 1.7 - fl(delta * 3)
    want_g[1][1, 3, 1] = T(T(1.7) - T(delta * T(3)))
    want_dc[1][1, 3, 1] = 4
    # D: clamp (reference mode: f - eps in T)
    want_g[1][2, 3, 0] = T(T(5.0) - T(eb))
    want_dc[1][2, 3, 0] = 5
    for k in range(2):
        assert go[k].tobytes() == want_g[k].tobytes(), k
        assert do[k].tobytes() == want_dc[k].tobytes(), k


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sync_kat_other_direction(dtype):
    T = np.dtype(dtype).type
    shape, f, g = _two_blocks_x(dtype)
    eb = 0.35
    delta = T(eb / 5)
    fb, gb = O.blocks(f, g, (2, 1, 1))
    dcb = [np.zeros_like(b, dtype=np.uint8) for b in fb]
    dcb[1][1, 2, 1] = 2      # block 1's outermost core toward -x: global (2, 1, 0) = G
    go, do = _sync_both(shape, eb, (2, 1, 1), fb, gb, dcb, [0, 1])
    b = T(T(2.3) - T(delta * T(2)))
    assert b != T(T(T(2.3) - delta) - delta) or dtype == np.float32
    assert go[0][1, 2, 3] == b and do[0][1, 2, 3] == 2
    assert go[1].tobytes() == gb[1].tobytes() and do[1].tobytes() == dcb[1].tobytes()
    rest = go[0].copy()
    rest[1, 2, 3] = gb[0][1, 2, 3]
    assert rest.tobytes() == gb[0].tobytes()


@pytest.mark.parametrize("policy", [O.GHOST_REFERENCE, O.GHOST_FULL])
def test_sync_kat_yz_edge_and_corner(policy):
    # the delta-counter exchange covers directions 0..11 under both policies
    # (pMSz.cu:1137), so the (0, 1, 1) edge is synchronised even where the initial
    # exchange dropped it; the stencil corners only under the full policy
    shape = (4, 4, 4)
    f = np.full(shape, 3.0)
    g = np.full(shape, 2.5)
    sub = (2, 2, 2)
    fb, gb = O.blocks(f, g, sub, ghost_policy=O.GHOST_FULL)  # every ghost holds a value
    dcb = [np.zeros_like(b, dtype=np.uint8) for b in fb]
    send = [0] * 8
    # block (0,0,0) = 0: core cell padded (1, 2, 2) = global (0, 1, 1), toward (0, 1, 1)
    dcb[0][2, 2, 1] = 2
    send[0] = 1
    # block (1,0,0) = 1: core corner padded (1, 2, 2) = global (2, 1, 1), toward (-1, 1, 1)
    dcb[1][2, 2, 1] = 3
    send[1] = 1
    go, do = _sync_both(shape, 0.5, sub, fb, gb, dcb, send, policy=policy)
    d = 0.5 / 5
    # block (0,1,1) = 6 receives global (0, 1, 1) in its ghost padded (1, 0, 0)
    assert do[6][0, 0, 1] == 2 and go[6][0, 0, 1] == 2.5 - d * 2
    # block (0,1,1) receives global (2, 1, 1) in its ghost padded (3, 0, 0): corner
    if policy == O.GHOST_FULL:
        assert do[6][0, 0, 3] == 3 and go[6][0, 0, 3] == 2.5 - d * 3
    else:
        assert do[6][0, 0, 3] == 0 and go[6][0, 0, 3] == 2.5


# ---------------------------------------------------------------- loops, hand-worked
def _ramp_case():
    """f = 0..5 along x (one row), g0 raises x = 3 to 4.25; eb = 1.25, q = 5,
    delta = 0.25 (exact).  Split 2 x 1 x 1: block 0 holds x = 0..2, block 1 x = 3..5."""
    f = np.arange(6, dtype=np.float64).reshape(1, 1, 6)
    g0 = f.copy()
    g0[0, 0, 3] = 4.25
    return f, g0, 1.25


def test_loop_kat_lazy_two_blocks():
    # Block 0 (x 0..2, ghost x 3 = 4.25): x = 2 ascends to its ghost as in f; no
    # violation: one sweep, ite = 1 + 1 = 2.
    # Block 1 (ghost x 2 = 2, x 3..5): x = 3 is a false maximum (4.25 above 2 and 4)
    # -> target x = 3; x = 4 is a false minimum (4 below 4.25 and 5) -> target its
    # original descent neighbour x = 3.  One edit: 4.25 - 0.25 = 4.0 (|3 - 4| <= eb),
    # dc = 1, send_flag (x = 3 is padded x 1).  Sweep 2: x = 3 (4.0) ties x = 4 (4.0),
    # the higher index ascends, the lower descends: no violation.  ite = 1 + 2 = 3.
    # Sync: block 1 sends [dc(x=3) = 1, dc(ghost x=2) = 0] toward -x; block 0's
    # ghost x = 3 takes dc 1 and 4.25 - 0.25 * 1 = 4.0.  Checks find nothing.
    f, g0, eb = _ramp_case()
    want = g0.copy()
    want[0, 0, 3] = 4.0
    for g, dc, st in (_oracle(f, g0, eb, O.LAZY), S.correct(f, g0, eb, (2, 1, 1), O.LAZY)):
        assert g.tobytes() == want.tobytes()
        assert dc.ravel().tolist() == [0, 0, 0, 1, 0, 0]
        assert (st["edit_sweeps"], st["check_sweeps"], st["outer_iters"]) == (3, 2, 1)
        assert st["ite"] == 2 and st["total_targets"] == 1


def test_loop_kat_sync_two_blocks():
    # Per outer iteration one edit sweep per block (ite += 2): block 0 finds nothing,
    # block 1 edits x = 3 once; the sync and the checks are as in the lazy case, so
    # one outer iteration: 2 edit sweeps, 2 checks, ite = 2.
    f, g0, eb = _ramp_case()
    want = g0.copy()
    want[0, 0, 3] = 4.0
    for g, dc, st in (_oracle(f, g0, eb, O.SYNC), S.correct(f, g0, eb, (2, 1, 1), O.SYNC)):
        assert g.tobytes() == want.tobytes()
        assert (st["edit_sweeps"], st["check_sweeps"], st["outer_iters"]) == (2, 2, 1)
        assert st["ite"] == 2 and st["total_targets"] == 1


def test_loop_kat_lazy_shared_cell_two_outer_iterations():
    # f = 0..5, g0 = [0, 1.5, 1.75, 4, 4, 5], eb = 1.25, delta = 0.25.
    # Block 0 (x 0..2; ghost x 3 = 4): x = 2 (1.75) ascends to the ghost and descends
    # to x = 1 (1.5): no violation.
    # Block 1 (ghost x 2 = 1.75; x 3..5): x = 3 (4) ties x = 4 (4): the higher index
    # x = 4 ascends, as in f; x = 4 descends to x = 3 (lower index), as in f.  No
    # violation either.  Nothing is edited: one outer iteration with one sweep per
    # block, ite = 2.
    f = np.arange(6, dtype=np.float64).reshape(1, 1, 6)
    g0 = np.array([0, 1.5, 1.75, 4, 4, 5], np.float64).reshape(1, 1, 6)
    for g, dc, st in (_oracle(f, g0, 1.25, O.LAZY), S.correct(f, g0, 1.25, (2, 1, 1), O.LAZY)):
        assert g.tobytes() == g0.tobytes() and not dc.any()
        assert (st["edit_sweeps"], st["outer_iters"], st["ite"]) == (2, 1, 2)


def _oracle(f, g0, eb, mode, sub=(2, 1, 1), **kw):
    g, dc, st, _, rc = O.correct(f, g0, eb, sub=sub, mode=mode, **kw)
    assert rc == 0
    return g, dc, st


# ---------------------------------------------------------------- loops, restatement
CASES = [((6, 5, 7), (2, 2, 2)), ((5, 6, 4), (3, 2, 1)), ((7, 4, 5), (1, 2, 3)),
         ((4, 9, 6), (2, 3, 1)), ((6, 6, 6), (2, 1, 2))]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("policy", [O.GHOST_REFERENCE, O.GHOST_FULL])
@pytest.mark.parametrize("mode", [O.LAZY, O.SYNC])
def test_restatement_matches_oracle(case, policy, mode):
    shape, sub = CASES[case]
    for dtype, alphabet, clamp in ((np.float64, None, O.CLAMP_REFERENCE),
                                   (np.float32, 6, O.CLAMP_BOUND_SAFE)):
        rng = np.random.default_rng(100 * case + 7 * mode + policy)
        if alphabet:
            f = rng.integers(0, alphabet, size=shape).astype(dtype)
            eb = 0.8
        else:
            f = rng.standard_normal(shape).astype(dtype)
            eb = 0.4
        g0 = (f + rng.uniform(-eb, eb, size=shape) * 0.9).astype(dtype)
        if alphabet:   # ties in g0 as well
            g0 = np.round(g0 * 4) / 4
            g0 = np.where(np.abs(g0 - f) <= eb * 0.9, g0, f).astype(dtype)
        g, dc, st, _, rc = O.correct(f, g0, eb, sub=sub, mode=mode, ghost_policy=policy,
                                     clamp_mode=clamp)
        assert rc == 0
        rg, rdc, rst = S.correct(f, g0, eb, sub, mode=mode, policy=policy, clamp_mode=clamp)
        assert g.tobytes() == rg.tobytes(), (dtype, shape, sub)
        assert dc.tobytes() == rdc.tobytes()
        for k in ("edit_sweeps", "check_sweeps", "outer_iters", "ite", "total_targets"):
            assert st[k] == rst[k], (k, st[k], rst[k])
        assert st["total_targets"] > 0
