"""ctypes view of the CPU oracle (libpmsz_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the CPU-baseline leg of ``bench.py`` as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpmsz_oracle.so")

LAZY, SYNC, NAIVE = 0, 1, 2
CLAMP_REFERENCE, CLAMP_BOUND_SAFE = 0, 1
OK, EINVAL, ENOMEM = 0, -1, -2
GHOST_REFERENCE, GHOST_FULL = 0, 1
NOT_CONVERGED = -3


class Params(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int),
        ("nx", ctypes.c_int64),
        ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64),
        ("abs_eb", ctypes.c_double),
        ("q", ctypes.c_int),
        ("sub", ctypes.c_int * 3),
        ("mode", ctypes.c_int),
        ("max_sweeps", ctypes.c_int),
        ("clamp_mode", ctypes.c_int),
        ("ghost_policy", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("edit_sweeps", ctypes.c_int64),
        ("check_sweeps", ctypes.c_int64),
        ("outer_iters", ctypes.c_int64),
        ("ite", ctypes.c_int64),
        ("edited_voxels", ctypes.c_int64),
        ("clamped_voxels", ctypes.c_int64),
        ("total_targets", ctypes.c_int64),
        ("final_cnt", ctypes.c_uint32 * 3),
        ("status", ctypes.c_int),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "final_cnt"}
        d["final_cnt"] = list(self.final_cnt)
        return d


class Trace(ctypes.Structure):
    _fields_ = [
        ("cap", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("cnt", ctypes.POINTER(ctypes.c_uint32)),
        ("targets", ctypes.POINTER(ctypes.c_int64)),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.orc_codes.argtypes = [i32, i64, i64, i64, vp, vp]
        L.orc_labels.argtypes = [i32, i64, i64, i64, vp, vp, vp]
        L.orc_sweep.argtypes = [ctypes.POINTER(Params), vp, vp, vp, i32, vp, vp, vp]
        L.orc_correct.argtypes = [ctypes.POINTER(Params), vp, vp, vp, ctypes.POINTER(Stats),
                                  ctypes.POINTER(Trace)]
        L.orc_edit_symbols.argtypes = [ctypes.POINTER(Params), vp, vp, vp, vp, vp]
        L.orc_codec_dq.argtypes = [ctypes.POINTER(Params), vp, vp, vp]
        L.orc_bits_for_max.argtypes = [ctypes.c_uint]
        L.orc_bits_for_max.restype = ctypes.c_uint
        L.orc_encode_edits.argtypes = [ctypes.POINTER(Params), vp, vp, vp, vp, vp,
                                       ctypes.POINTER(EditBlock), vp]
        L.orc_within.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_within.restype = i32
        L.orc_split.argtypes = [i64, i32, i32, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.orc_split.restype = None
        L.orc_dq_compress.argtypes = [ctypes.POINTER(Params), vp, vp, vp, i64,
                                      ctypes.POINTER(i64)]
        L.orc_dq_decompress.argtypes = [vp, i64, vp]
        L.orc_blocks.argtypes = [ctypes.POINTER(Params), vp, vp, vp, vp]
        L.orc_sync_blocks.argtypes = [ctypes.POINTER(Params), vp, vp, vp, vp]
        _lib = L
    return _lib


def _dtype_code(a):
    if a.dtype == np.float32:
        return 0
    if a.dtype == np.float64:
        return 1
    raise TypeError(a.dtype)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_params(shape, dtype, eb, q=5, sub=(1, 1, 1), mode=LAZY, max_sweeps=100000,
                clamp_mode=CLAMP_BOUND_SAFE, ghost_policy=GHOST_FULL):
    """shape is (nz, ny, nx) -- numpy C order with x fastest."""
    nz, ny, nx = shape
    p = Params()
    p.dtype = 0 if np.dtype(dtype) == np.float32 else 1
    p.nx, p.ny, p.nz = nx, ny, nz
    p.abs_eb = float(eb)
    p.q = q
    p.sub[0], p.sub[1], p.sub[2] = sub
    p.mode = mode
    p.max_sweeps = max_sweeps
    p.clamp_mode = clamp_mode
    p.ghost_policy = ghost_policy
    return p


def codes(field):
    field = np.ascontiguousarray(field)
    nz, ny, nx = field.shape
    out = np.empty(field.shape, np.uint8)
    rc = lib().orc_codes(_dtype_code(field), nx, ny, nz, _ptr(field), _ptr(out))
    assert rc == 0, rc
    return out


def labels(field):
    field = np.ascontiguousarray(field)
    nz, ny, nx = field.shape
    asc = np.empty(field.shape, np.int64)
    desc = np.empty(field.shape, np.int64)
    rc = lib().orc_labels(_dtype_code(field), nx, ny, nz, _ptr(field), _ptr(asc), _ptr(desc))
    assert rc == 0, rc
    return asc, desc


def sweep(f, g, dc, eb, edit=True, **kw):
    """One snapshot sweep on a single domain; returns (g', dc', counts, ntargets, mask)."""
    f = np.ascontiguousarray(f)
    g = np.array(g, dtype=f.dtype, copy=True)
    dc = np.array(dc, dtype=np.uint8, copy=True)
    p = make_params(f.shape, f.dtype, eb, **kw)
    counts = np.zeros(3, np.uint32)
    nt = ctypes.c_int64(0)
    mask = np.zeros(f.shape, np.uint8)
    rc = lib().orc_sweep(ctypes.byref(p), _ptr(f), _ptr(g), _ptr(dc), 1 if edit else 0,
                         _ptr(counts), ctypes.cast(ctypes.byref(nt), ctypes.c_void_p), _ptr(mask))
    assert rc == 0, rc
    return g, dc, counts, nt.value, mask


def correct(f, g0, eb, trace_cap=0, **kw):
    """Full correction; returns (g, dc, stats dict, trace or None, rc)."""
    f = np.ascontiguousarray(f)
    g = np.array(g0, dtype=f.dtype, copy=True)
    dc = np.zeros(f.shape, np.uint8)
    p = make_params(f.shape, f.dtype, eb, **kw)
    st = Stats()
    tr = None
    trp = ctypes.POINTER(Trace)()
    if trace_cap:
        cnt = np.zeros(3 * trace_cap, np.uint32)
        tg = np.zeros(trace_cap, np.int64)
        tr = Trace(trace_cap, 0, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                   tg.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        trp = ctypes.pointer(tr)
    rc = lib().orc_correct(ctypes.byref(p), _ptr(f), _ptr(g), _ptr(dc), ctypes.byref(st), trp)
    trace = None
    if tr is not None:
        trace = {"cnt": cnt[: 3 * tr.n].reshape(-1, 3).copy(), "targets": tg[: tr.n].copy()}
    return g, dc, st.as_dict(), trace, rc


def edit_symbols(f, g0, g, eb, **kw):
    p = make_params(f.shape, f.dtype, eb, **kw)
    sym = np.zeros(f.shape, np.uint8)
    ed = ctypes.c_int64(0)
    rc = lib().orc_edit_symbols(ctypes.byref(p), _ptr(f), _ptr(g0), _ptr(g), _ptr(sym),
                                ctypes.cast(ctypes.byref(ed), ctypes.c_void_p))
    assert rc == 0, rc
    return sym, ed.value


class EditBlock(ctypes.Structure):
    _fields_ = [("voxels", ctypes.c_int64), ("packed_offset", ctypes.c_int64),
                ("packed_bytes", ctypes.c_int64), ("value_offset", ctypes.c_int64),
                ("values", ctypes.c_int64)]


def bits_for_max(q):
    return int(lib().orc_bits_for_max(int(q)))


def encode_edits(f, g0, g, eb, **kw):
    """(packed bytes, clamp values, blocks, edited, rc) of the edit side channel:
    per rank block in rank order, as pmsz_encode_edits lays them out."""
    p = make_params(f.shape, f.dtype, eb, **kw)
    nb = p.sub[0] * p.sub[1] * p.sub[2]
    nbits = bits_for_max(p.q)
    packed = np.zeros(f.size * nbits // 8 + 32 * nb, np.uint8)
    vals = np.zeros(f.size, f.dtype)
    blocks = (EditBlock * nb)()
    ed = ctypes.c_int64(0)
    rc = lib().orc_encode_edits(ctypes.byref(p), _ptr(f), _ptr(g0), _ptr(g), _ptr(packed),
                                _ptr(vals), blocks,
                                ctypes.cast(ctypes.byref(ed), ctypes.c_void_p))
    bl = [{k: getattr(b, k) for k, _ in EditBlock._fields_} for b in blocks]
    nv = sum(b["values"] for b in bl)
    end = bl[-1]["packed_offset"] + bl[-1]["packed_bytes"]
    return packed[:end], vals[:nv], bl, ed.value, rc


DQ_RADIUS = 512
DQ_BINS = 2 * DQ_RADIUS


def codec_dq(f, eb, sub=(1, 1, 1)):
    """(g0, hist) of the dual-quantisation Lorenzo codec: hist[b] holds block b's
    DQ_BINS residual bins, then its escapes and verbatim voxels."""
    p = make_params(f.shape, f.dtype, eb, sub=sub)
    nb = sub[0] * sub[1] * sub[2]
    g0 = np.empty_like(f)
    hist = np.zeros((nb, DQ_BINS + 2), np.uint64)
    rc = lib().orc_codec_dq(ctypes.byref(p), _ptr(f), _ptr(g0), _ptr(hist))
    assert rc == 0, rc
    return g0, hist


def dq_compress(f, eb, sub=(1, 1, 1)):
    """(g0, stream) of the codec's compressed stream: stream is a uint8 array."""
    f = np.ascontiguousarray(f)
    p = make_params(f.shape, f.dtype, eb, sub=sub)
    g0 = np.empty_like(f)
    need = ctypes.c_int64(0)
    lib().orc_dq_compress(ctypes.byref(p), _ptr(f), _ptr(g0), None, 0, ctypes.byref(need))
    buf = np.zeros(max(need.value, 64), np.uint8)
    rc = lib().orc_dq_compress(ctypes.byref(p), _ptr(f), _ptr(g0), _ptr(buf), buf.size,
                               ctypes.byref(need))
    assert rc == 0, rc
    return g0, buf[:need.value]


def dq_stream_shape(stream):
    """(shape (nz, ny, nx), dtype, sub, abs_eb) from a stream's header."""
    import struct
    _, dt, nx, ny, nz, s0, s1, s2, nb, eb = struct.unpack_from("<IIqqqiiiId", bytes(stream[:56]))
    return (nz, ny, nx), (np.float32 if dt == 0 else np.float64), (s0, s1, s2), eb


def dq_decompress(stream):
    stream = np.ascontiguousarray(stream, dtype=np.uint8)
    shape, dt, _, _ = dq_stream_shape(stream)
    g0 = np.empty(shape, dt)
    rc = lib().orc_dq_decompress(_ptr(stream), stream.size, _ptr(g0))
    assert rc == 0, rc
    return g0


def clamp(fv, eb, clamp_mode, dtype=np.float32):
    """Clamp value of the edit primitive for original value fv (a float of dtype)."""
    L = lib()
    fn = L.orc_clamp_f32 if np.dtype(dtype) == np.float32 else L.orc_clamp_f64
    fn.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
    fn.restype = ctypes.c_double
    return np.dtype(dtype).type(fn(float(fv), float(eb), int(clamp_mode)))


def within(a, b, eb):
    return bool(lib().orc_within(float(a), float(b), float(eb)))


def split(n, parts, coord):
    lo, off = ctypes.c_int64(0), ctypes.c_int64(0)
    lib().orc_split(n, parts, coord, ctypes.byref(lo), ctypes.byref(off))
    return lo.value, off.value


def block_shapes(shape, sub):
    """Padded (lz+2, ly+2, lx+2) shapes of the subdomain blocks, in the order cx
    fastest, then cy, then cz, with their (oz, oy, ox) core offsets."""
    nz, ny, nx = shape
    out = []
    for cz in range(sub[2]):
        for cy in range(sub[1]):
            for cx in range(sub[0]):
                lx, ox = split(nx, sub[0], cx)
                ly, oy = split(ny, sub[1], cy)
                lz, oz = split(nz, sub[2], cz)
                out.append(((lz + 2, ly + 2, lx + 2), (oz, oy, ox)))
    return out


def blocks(f, g, sub, ghost_policy=GHOST_FULL, eb=1.0):
    """Padded blocks (f, g) of every subdomain after set-up: the cores and the
    initial ghost values.  Returns two lists of (lz+2, ly+2, lx+2) arrays."""
    p = make_params(f.shape, f.dtype, eb, sub=sub, ghost_policy=ghost_policy)
    shp = block_shapes(f.shape, sub)
    tot = sum(int(np.prod(s)) for s, _ in shp)
    fo = np.empty(tot, f.dtype)
    go = np.empty(tot, f.dtype)
    rc = lib().orc_blocks(ctypes.byref(p), _ptr(np.ascontiguousarray(f)),
                          _ptr(np.ascontiguousarray(g)), _ptr(fo), _ptr(go))
    assert rc == 0, rc
    fb, gb, off = [], [], 0
    for s, _ in shp:
        n = int(np.prod(s))
        fb.append(fo[off:off + n].reshape(s).copy())
        gb.append(go[off:off + n].reshape(s).copy())
        off += n
    return fb, gb


def sync_blocks(shape, eb, sub, fb, gb, dcb, send, q=5, ghost_policy=GHOST_FULL,
                clamp_mode=CLAMP_BOUND_SAFE):
    """One delta-counter ghost sync on the padded blocks (lists as from blocks(),
    dc blocks uint8); returns the new (g blocks, dc blocks)."""
    dtype = fb[0].dtype
    p = make_params(shape, dtype, eb, q=q, sub=sub, ghost_policy=ghost_policy,
                    clamp_mode=clamp_mode)
    fcat = np.concatenate([b.ravel() for b in fb]).astype(dtype)
    gcat = np.concatenate([b.ravel() for b in gb]).astype(dtype)
    dcat = np.concatenate([b.ravel() for b in dcb]).astype(np.uint8)
    flags = np.ascontiguousarray(send, dtype=np.int32)
    rc = lib().orc_sync_blocks(ctypes.byref(p), _ptr(fcat), _ptr(gcat), _ptr(dcat), _ptr(flags))
    assert rc == 0, rc
    go, do, off = [], [], 0
    for b in fb:
        n = b.size
        go.append(gcat[off:off + n].reshape(b.shape).copy())
        do.append(dcat[off:off + n].reshape(b.shape).copy())
        off += n
    return go, do
