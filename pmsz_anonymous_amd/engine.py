"""Python view of libpmsz.so (include/pmsz.h).

The host-side mirror of the reference's correction entry point: ``correct`` takes
the original field ``f`` and the decompressed field ``g`` (torch tensors on the
GPU, or numpy arrays on the host) with an absolute error bound, the reference's
``q``, subdomain grid and sync mode (c_loops, pMSz.cu:1872-2074), and returns the
corrected field with the run's statistics.  Errors raise ``PmszError`` carrying
the C-ABI code.  There is no CPU fallback: without the built library every call
raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpmsz.so")

F32, F64 = 0, 1
LAZY, SYNC, NAIVE = 0, 1, 2
CLAMP_REFERENCE, CLAMP_BOUND_SAFE = 0, 1
GHOST_REFERENCE, GHOST_FULL = 0, 1

OK, EINVAL, ENOMEM, NOT_CONVERGED, EDEVICE, ETOOBIG, EBUSY = 0, -1, -2, -3, -4, -5, -6

EXPORTED = [
    "pmsz_default_params", "pmsz_correct", "pmsz_correct_batch", "pmsz_correct_host",
    "pmsz_ctx_create", "pmsz_ctx_destroy", "pmsz_ctx_correct", "pmsz_ctx_correct_batch",
    "pmsz_ctx_bytes", "pmsz_ctx_trim", "pmsz_workspace_bytes",
    "pmsz_codes", "pmsz_labels", "pmsz_edit_symbols", "pmsz_edit_packed_bytes",
    "pmsz_encode_edits", "pmsz_codec_dq", "pmsz_time_first_pass", "pmsz_strerror",
    "pmsz_version", "pmsz_default_tuning", "pmsz_ctx_set_tuning", "pmsz_ctx_get_tuning",
    "pmsz_autotune", "pmsz_test_step_budget",
]


class PmszError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().pmsz_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


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
        ("touched_voxels", ctypes.c_int64),
        ("clamped_voxels", ctypes.c_int64),
        ("total_targets", ctypes.c_int64),
        ("final_cnt", ctypes.c_uint32 * 3),
        ("status", ctypes.c_int),
        ("dense_passes", ctypes.c_int64),
        ("sparse_voxels", ctypes.c_int64),
        ("t_total_s", ctypes.c_double),
        ("t_dense_s", ctypes.c_double),
        ("t_label_s", ctypes.c_double),
        ("bytes_algorithmic", ctypes.c_double),
        ("trace_n", ctypes.c_int64),
        ("trace_cnt", (ctypes.c_uint32 * 3) * 64),
        ("trace_targets", ctypes.c_uint32 * 64),
    ]

    def as_dict(self):
        skip = ("final_cnt", "trace_cnt", "trace_targets")
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in skip}
        d["final_cnt"] = list(self.final_cnt)
        n = int(self.trace_n)
        d["trace_cnt"] = [list(self.trace_cnt[i]) for i in range(n)]
        d["trace_targets"] = [int(self.trace_targets[i]) for i in range(n)]
        return d


_lib = None


class Tuning(ctypes.Structure):
    """pmsz_tuning: schedule and tile parameters (speed only, never results)."""
    _fields_ = [("dense_div", ctypes.c_int), ("tail_targets", ctypes.c_int),
                ("tail_blocks", ctypes.c_int), ("solo_targets", ctypes.c_int),
                ("zchunk", ctypes.c_int), ("graphs", ctypes.c_int),
                ("stream_div", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EditBlock(ctypes.Structure):
    """pmsz_edit_block: one rank block of the edit side channel."""
    _fields_ = [("voxels", ctypes.c_int64), ("packed_offset", ctypes.c_int64),
                ("packed_bytes", ctypes.c_int64), ("value_offset", ctypes.c_int64),
                ("values", ctypes.c_int64)]


class EditInfo(ctypes.Structure):
    """pmsz_edit_info: totals of pmsz_encode_edits."""
    _fields_ = [("nbits", ctypes.c_int32), ("blocks", ctypes.c_int32),
                ("edited", ctypes.c_int64), ("values", ctypes.c_int64),
                ("packed_capacity", ctypes.c_int64)]


def _tuning(d):
    t = Tuning()
    lib().pmsz_default_tuning(ctypes.byref(t))
    for k, v in (d or {}).items():
        setattr(t, k, int(v))
    return t


def set_tuning(tuning, ctx=None):
    """Set the tuning (a dict of Tuning fields; missing ones take the defaults) of
    a context (None: the default context)."""
    _check(lib().pmsz_ctx_set_tuning(ctx._h if ctx else None, ctypes.byref(_tuning(tuning))),
           "pmsz_ctx_set_tuning")


def get_tuning(ctx=None):
    t = Tuning()
    _check(lib().pmsz_ctx_get_tuning(ctx._h if ctx else None, ctypes.byref(t)),
           "pmsz_ctx_get_tuning")
    return t.as_dict()


def autotune(fs, g0s, ebs, *, reps=2, ctx=None, per_field=None, **kw):
    """Autotune on copies of CUDA tensors g0s[k] against fs[k] (one field, or a
    batch): the fastest tuning is kept in the context and returned as
    (dict, seconds).  g0s are not modified."""
    import torch
    n, params, pf, pg, _ = _batch_args(fs, g0s, ebs, None, per_field, kw)
    best = Tuning()
    secs = ctypes.c_double()
    torch.cuda.current_stream().synchronize()
    _check(lib().pmsz_autotune(ctx._h if ctx else None, n, params, pf, pg, reps,
                               ctypes.byref(best), ctypes.byref(secs)), "pmsz_autotune")
    return best.as_dict(), secs.value


def lib_path():
    return _LIB_PATH


def lib():
    """The loaded libpmsz.so; raises if it was not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built: run python -m pmsz_anonymous_amd.build")
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        P = ctypes.POINTER(Params)
        S = ctypes.POINTER(Stats)
        L.pmsz_default_params.argtypes = [P]
        L.pmsz_default_params.restype = None
        L.pmsz_correct.argtypes = [P, vp, vp, vp, S]
        L.pmsz_correct_host.argtypes = [P, vp, vp, vp, S]
        L.pmsz_correct_batch.argtypes = [i32, P, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                         ctypes.POINTER(vp), S]
        L.pmsz_ctx_create.argtypes = [ctypes.POINTER(vp)]
        L.pmsz_ctx_destroy.argtypes = [vp]
        L.pmsz_ctx_correct.argtypes = [vp, P, vp, vp, vp, S]
        L.pmsz_ctx_correct_batch.argtypes = [vp, i32, P, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                             ctypes.POINTER(vp), S]
        L.pmsz_ctx_trim.argtypes = [vp]
        L.pmsz_ctx_bytes.argtypes = [vp]
        L.pmsz_ctx_bytes.restype = i64
        L.pmsz_workspace_bytes.argtypes = [i32, P]
        L.pmsz_workspace_bytes.restype = i64
        L.pmsz_codes.argtypes = [i32, i64, i64, i64, vp, vp]
        L.pmsz_labels.argtypes = [i32, i64, i64, i64, vp, vp, vp]
        L.pmsz_edit_symbols.argtypes = [P, vp, vp, vp, vp, ctypes.POINTER(i64)]
        L.pmsz_edit_packed_bytes.argtypes = [P, ctypes.POINTER(i64)]
        L.pmsz_codec_dq.argtypes = [P, vp, vp, vp, ctypes.POINTER(i64 * 3)]
        L.pmsz_encode_edits.argtypes = [P, vp, vp, vp, vp, vp, ctypes.POINTER(EditBlock),
                                        ctypes.POINTER(EditInfo)]
        L.pmsz_codec_dq_stream_bound.argtypes = [P, ctypes.POINTER(i64)]
        L.pmsz_codec_dq_compress.argtypes = [P, vp, vp, vp, i64, ctypes.POINTER(i64),
                                             ctypes.POINTER(i64 * 3)]
        L.pmsz_codec_dq_stream_params.argtypes = [vp, i64, P]
        L.pmsz_codec_dq_decompress.argtypes = [vp, i64, vp]
        L.pmsz_encode_edits_noalloc.argtypes = [P, vp, vp, vp, vp, vp, ctypes.POINTER(EditBlock),
                                                ctypes.POINTER(EditInfo), vp]
        L.pmsz_time_first_pass.argtypes = [P, vp, vp, i32, ctypes.POINTER(ctypes.c_double)]
        L.pmsz_test_step_budget.argtypes = [i32, i64, vp, vp, vp, ctypes.c_double, i32, i32, vp]
        L.pmsz_strerror.argtypes = [i32]
        L.pmsz_strerror.restype = ctypes.c_char_p
        L.pmsz_version.argtypes = []
        L.pmsz_version.restype = ctypes.c_char_p
        T = ctypes.POINTER(Tuning)
        L.pmsz_default_tuning.argtypes = [T]
        L.pmsz_default_tuning.restype = None
        L.pmsz_ctx_set_tuning.argtypes = [vp, T]
        L.pmsz_ctx_get_tuning.argtypes = [vp, T]
        L.pmsz_autotune.argtypes = [vp, i32, P, ctypes.POINTER(vp), ctypes.POINTER(vp), i32, T,
                                    ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _dtype_code(dt):
    s = str(dt)
    if s in ("float32", "torch.float32"):
        return F32
    if s in ("float64", "torch.float64"):
        return F64
    raise TypeError(f"unsupported dtype {dt}")


def make_params(shape, dtype, abs_eb, q=5, sub=(1, 1, 1), mode=LAZY, max_sweeps=100000,
                clamp_mode=CLAMP_BOUND_SAFE, ghost_policy=GHOST_FULL):
    """shape = (nz, ny, nx), C order with x fastest."""
    p = Params()
    lib().pmsz_default_params(ctypes.byref(p))
    nz, ny, nx = (int(s) for s in shape)
    p.dtype = _dtype_code(dtype)
    p.nx, p.ny, p.nz = nx, ny, nz
    p.abs_eb = float(abs_eb)
    p.q = int(q)
    p.sub[0], p.sub[1], p.sub[2] = (int(s) for s in sub)
    p.mode = int(mode)
    p.max_sweeps = int(max_sweeps)
    p.clamp_mode = int(clamp_mode)
    p.ghost_policy = int(ghost_policy)
    return p


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _check(rc, what, allow=()):
    if rc != OK and rc not in allow:
        raise PmszError(rc, what)
    return rc


def correct(f, g, abs_eb, *, dc=None, raise_not_converged=True, **kw):
    """Correct ``g`` in place (torch CUDA tensors) or in a copy (numpy arrays).

    Returns ``(g, stats)``; ``stats['status']`` is NOT_CONVERGED (-3) when the sweep
    cap was hit and ``raise_not_converged`` is False.
    """
    p = make_params(f.shape, f.dtype, abs_eb, **kw)
    st = Stats()
    allow = () if raise_not_converged else (NOT_CONVERGED,)
    if _is_torch(f):
        import torch
        if not (f.is_cuda and g.is_cuda):
            raise ValueError("torch inputs must be on the GPU")
        if f.shape != g.shape or f.dtype != g.dtype:
            raise ValueError("f and g must match in shape and dtype")
        f = f.contiguous()
        if not g.is_contiguous():
            raise ValueError("g must be contiguous (it is corrected in place)")
        dptr = None
        if dc is not None:
            if dc.dtype != torch.uint8 or dc.numel() != f.numel() or not dc.is_cuda:
                raise ValueError("dc must be a uint8 CUDA tensor with one entry per voxel")
            dptr = ctypes.c_void_p(dc.data_ptr())
        torch.cuda.current_stream().synchronize()
        rc = lib().pmsz_correct(ctypes.byref(p), ctypes.c_void_p(f.data_ptr()),
                                ctypes.c_void_p(g.data_ptr()), dptr, ctypes.byref(st))
        _check(rc, "pmsz_correct", allow)
        return g, st.as_dict()
    f = np.ascontiguousarray(f)
    out = np.array(g, dtype=f.dtype, copy=True, order="C")
    dptr = None
    if dc is not None:
        if dc.dtype != np.uint8 or dc.size != f.size or not dc.flags.c_contiguous:
            raise ValueError("dc must be a contiguous uint8 array with one entry per voxel")
        dptr = dc.ctypes.data_as(ctypes.c_void_p)
    rc = lib().pmsz_correct_host(ctypes.byref(p), f.ctypes.data_as(ctypes.c_void_p),
                                 out.ctypes.data_as(ctypes.c_void_p), dptr, ctypes.byref(st))
    _check(rc, "pmsz_correct_host", allow)
    return out, st.as_dict()


def _batch_args(fs, gs, ebs, dcs, per_field, kw):
    import torch
    n = len(fs)
    if not (len(gs) == n and len(ebs) == n) or n == 0:
        raise ValueError("fs, gs and ebs must be non-empty lists of one length")
    params = (Params * n)()
    for k in range(n):
        f, g = fs[k], gs[k]
        if not (f.is_cuda and g.is_cuda and f.is_contiguous() and g.is_contiguous()):
            raise ValueError("batch fields must be contiguous CUDA tensors")
        if f.shape != g.shape or f.dtype != g.dtype:
            raise ValueError("f and g must match in shape and dtype")
        extra = dict(kw)
        if per_field is not None:
            extra.update(per_field[k])
        params[k] = make_params(f.shape, f.dtype, ebs[k], **extra)
    vpa = ctypes.c_void_p * n
    pf = vpa(*[f.data_ptr() for f in fs])
    pg = vpa(*[g.data_ptr() for g in gs])
    pd = None
    if dcs is not None:
        for k in range(n):
            if dcs[k] is not None and (dcs[k].dtype != torch.uint8 or dcs[k].numel() != fs[k].numel()):
                raise ValueError("dcs[k] must be a uint8 CUDA tensor with one entry per voxel")
        pd = vpa(*[(d.data_ptr() if d is not None else None) for d in dcs])
    return n, params, pf, pg, pd


def workspace_bytes(shapes_dtypes, **kw):
    """Library-owned device bytes a batch run of fields of these (shape, dtype)
    pairs needs (pmsz_workspace_bytes)."""
    n = len(shapes_dtypes)
    params = (Params * n)()
    for k, (shape, dt) in enumerate(shapes_dtypes):
        params[k] = make_params(shape, dt, 1.0, **kw)
    v = lib().pmsz_workspace_bytes(n, params)
    if v < 0:
        raise PmszError(int(v), "pmsz_workspace_bytes")
    return int(v)


class Context:
    """A correction context of its own (pmsz_ctx): its own HIP stream, scratch
    state and lock; runs in different contexts overlap on the device."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().pmsz_ctx_create(ctypes.byref(h)), "pmsz_ctx_create")
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            lib().pmsz_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes(self):
        return int(lib().pmsz_ctx_bytes(self._h))

    def trim(self):
        _check(lib().pmsz_ctx_trim(self._h), "pmsz_ctx_trim")

    def correct(self, f, g, abs_eb, *, dc=None, raise_not_converged=True, **kw):
        """As ``correct`` for CUDA tensors, in this context."""
        import torch
        p = make_params(f.shape, f.dtype, abs_eb, **kw)
        st = Stats()
        if not (f.is_cuda and g.is_cuda and f.is_contiguous() and g.is_contiguous()):
            raise ValueError("context inputs must be contiguous CUDA tensors")
        dptr = ctypes.c_void_p(dc.data_ptr()) if dc is not None else None
        torch.cuda.current_stream().synchronize()
        rc = lib().pmsz_ctx_correct(self._h, ctypes.byref(p), ctypes.c_void_p(f.data_ptr()),
                                    ctypes.c_void_p(g.data_ptr()), dptr, ctypes.byref(st))
        _check(rc, "pmsz_ctx_correct", () if raise_not_converged else (NOT_CONVERGED,))
        return g, st.as_dict()

    def correct_batch(self, fs, gs, ebs, *, dcs=None, per_field=None, raise_not_converged=True,
                      sync_torch=True, **kw):
        """As ``correct_batch``, in this context."""
        import torch
        n, params, pf, pg, pd = _batch_args(fs, gs, ebs, dcs, per_field, kw)
        stats = (Stats * n)()
        if sync_torch:
            torch.cuda.current_stream().synchronize()
        rc = lib().pmsz_ctx_correct_batch(self._h, n, params, pf, pg, pd, stats)
        _check(rc, "pmsz_ctx_correct_batch")
        out = [stats[k].as_dict() for k in range(n)]
        if raise_not_converged:
            for k, st in enumerate(out):
                if st["status"] != OK:
                    raise PmszError(st["status"], f"pmsz_ctx_correct_batch field {k}")
        return out


def correct_batch(fs, gs, ebs, *, dcs=None, per_field=None, raise_not_converged=True, **kw):
    """Correct a batch of CUDA tensors ``gs[k]`` in place against ``fs[k]`` with
    bound ``ebs[k]`` in one run (pmsz_correct_batch).  ``kw`` applies to every
    field (mode and max_sweeps must be common); ``per_field[k]`` (a dict of
    make_params keywords, e.g. q or clamp_mode) to field k.  Returns the list of
    per-field stats dicts."""
    import torch
    n, params, pf, pg, pd = _batch_args(fs, gs, ebs, dcs, per_field, kw)
    stats = (Stats * n)()
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_correct_batch(n, params, pf, pg, pd, stats)
    _check(rc, "pmsz_correct_batch")
    out = [stats[k].as_dict() for k in range(n)]
    if raise_not_converged:
        for k, st in enumerate(out):
            if st["status"] != OK:
                raise PmszError(st["status"], f"pmsz_correct_batch field {k}")
    return out


def codes(field):
    """Original classification codes of a CUDA tensor (uint8 tensor, same shape)."""
    import torch
    field = field.contiguous()
    nz, ny, nx = field.shape
    out = torch.empty(field.shape, dtype=torch.uint8, device=field.device)
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_codes(_dtype_code(field.dtype), nx, ny, nz, ctypes.c_void_p(field.data_ptr()),
                          ctypes.c_void_p(out.data_ptr()))
    _check(rc, "pmsz_codes")
    return out


def labels(field):
    """Ascending / descending Morse-Smale labels (int32 CUDA tensors)."""
    import torch
    field = field.contiguous()
    nz, ny, nx = field.shape
    asc = torch.empty(field.shape, dtype=torch.int32, device=field.device)
    desc = torch.empty(field.shape, dtype=torch.int32, device=field.device)
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_labels(_dtype_code(field.dtype), nx, ny, nz,
                           ctypes.c_void_p(field.data_ptr()), ctypes.c_void_p(asc.data_ptr()),
                           ctypes.c_void_p(desc.data_ptr()))
    _check(rc, "pmsz_labels")
    return asc, desc


def edit_symbols(f, g0, g, abs_eb, **kw):
    """(symbols uint8 tensor, edited count) of a correction (extract_edits_kernel)."""
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, **kw)
    sym = torch.empty(f.shape, dtype=torch.uint8, device=f.device)
    ed = ctypes.c_int64(0)
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_edit_symbols(ctypes.byref(p), ctypes.c_void_p(f.contiguous().data_ptr()),
                                 ctypes.c_void_p(g0.contiguous().data_ptr()),
                                 ctypes.c_void_p(g.contiguous().data_ptr()),
                                 ctypes.c_void_p(sym.data_ptr()), ctypes.byref(ed))
    _check(rc, "pmsz_edit_symbols")
    return sym, ed.value


def encode_edits(f, g0, g, abs_eb, **kw):
    """Edit side channel of a correction on CUDA tensors (pmsz_encode_edits).

    Returns (packed uint8 tensor, values tensor of f's dtype, blocks, info):
    blocks lists, per rank block in rank order, voxels / packed_offset /
    packed_bytes / value_offset / values; info holds nbits, blocks, edited,
    values and packed_capacity."""
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, **kw)
    cap = ctypes.c_int64(0)
    _check(lib().pmsz_edit_packed_bytes(ctypes.byref(p), ctypes.byref(cap)),
           "pmsz_edit_packed_bytes")
    nb = p.sub[0] * p.sub[1] * p.sub[2]
    packed = torch.empty(max(cap.value, 16), dtype=torch.uint8, device=f.device)
    vals = torch.empty(f.numel(), dtype=f.dtype, device=f.device)
    blocks = (EditBlock * nb)()
    info = EditInfo()
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_encode_edits(ctypes.byref(p), ctypes.c_void_p(f.contiguous().data_ptr()),
                                 ctypes.c_void_p(g0.contiguous().data_ptr()),
                                 ctypes.c_void_p(g.contiguous().data_ptr()),
                                 ctypes.c_void_p(packed.data_ptr()),
                                 ctypes.c_void_p(vals.data_ptr()), blocks, ctypes.byref(info))
    _check(rc, "pmsz_encode_edits")
    bl = [{k: getattr(b, k) for k, _ in EditBlock._fields_} for b in blocks]
    inf = {k: getattr(info, k) for k, _ in EditInfo._fields_}
    return packed[:cap.value], vals[:info.values], bl, inf


def encode_edits_noalloc(f, g0, g, abs_eb, packed=None, vals=None, **kw):
    """The edit side channel through pmsz_encode_edits_noalloc, which allocates no
    device memory: its scratch is the library's fixed per-device array.  packed /
    vals (nullable) are the caller's output tensors (packed: uint8, at least
    pmsz_edit_packed_bytes; vals: f's dtype, f.numel() elements).  Runs on torch's
    current stream.  Returns as encode_edits."""
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, **kw)
    cap = ctypes.c_int64(0)
    _check(lib().pmsz_edit_packed_bytes(ctypes.byref(p), ctypes.byref(cap)),
           "pmsz_edit_packed_bytes")
    nb = p.sub[0] * p.sub[1] * p.sub[2]
    if packed is None:
        packed = torch.empty(max(cap.value, 16), dtype=torch.uint8, device=f.device)
    if vals is None:
        vals = torch.empty(f.numel(), dtype=f.dtype, device=f.device)
    blocks = (EditBlock * nb)()
    info = EditInfo()
    stream = torch.cuda.current_stream(f.device).cuda_stream
    rc = lib().pmsz_encode_edits_noalloc(
        ctypes.byref(p), ctypes.c_void_p(f.contiguous().data_ptr()),
        ctypes.c_void_p(g0.contiguous().data_ptr()), ctypes.c_void_p(g.contiguous().data_ptr()),
        ctypes.c_void_p(packed.data_ptr()), ctypes.c_void_p(vals.data_ptr()), blocks,
        ctypes.byref(info), ctypes.c_void_p(stream))
    _check(rc, "pmsz_encode_edits_noalloc")
    bl = [{k: getattr(b, k) for k, _ in EditBlock._fields_} for b in blocks]
    inf = {k: getattr(info, k) for k, _ in EditInfo._fields_}
    return packed[:cap.value], vals[:info.values], bl, inf


DQ_SLOTS = 1026


def codec_dq(f, abs_eb, sub=(1, 1, 1)):
    """GPU dual-quantisation Lorenzo codec (pmsz_codec_dq) on a CUDA tensor f.
    Returns (g0 tensor, per-block histograms as a (blocks, 1026) uint64 array,
    info dict with compressed_bytes, escapes, verbatim)."""
    import numpy as np
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, sub=sub)
    nb = p.sub[0] * p.sub[1] * p.sub[2]
    g0 = torch.empty_like(f)
    hist = np.zeros((nb, DQ_SLOTS), np.uint64)
    info = (ctypes.c_int64 * 3)()
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_codec_dq(ctypes.byref(p), ctypes.c_void_p(f.contiguous().data_ptr()),
                             ctypes.c_void_p(g0.data_ptr()),
                             hist.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info))
    _check(rc, "pmsz_codec_dq")
    return g0, hist, {"compressed_bytes": info[0], "escapes": info[1], "verbatim": info[2]}


def codec_dq_compress(f, abs_eb, sub=(1, 1, 1)):
    """The codec's compressed stream on the device (pmsz_codec_dq_compress) for a
    CUDA tensor f.  Returns (g0 tensor, stream as a uint8 CUDA tensor, info dict
    with compressed_bytes (the entropy estimate), escapes, verbatim)."""
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, sub=sub)
    bound = ctypes.c_int64(0)
    _check(lib().pmsz_codec_dq_stream_bound(ctypes.byref(p), ctypes.byref(bound)),
           "pmsz_codec_dq_stream_bound")
    g0 = torch.empty_like(f)
    stream = torch.empty(bound.value, dtype=torch.uint8, device=f.device)
    nbytes = ctypes.c_int64(0)
    info = (ctypes.c_int64 * 3)()
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_codec_dq_compress(ctypes.byref(p), ctypes.c_void_p(f.contiguous().data_ptr()),
                                      ctypes.c_void_p(g0.data_ptr()),
                                      ctypes.c_void_p(stream.data_ptr()), bound.value,
                                      ctypes.byref(nbytes), ctypes.byref(info))
    _check(rc, "pmsz_codec_dq_compress")
    return g0, stream[:nbytes.value], {"compressed_bytes": info[0], "escapes": info[1],
                                       "verbatim": info[2]}


def codec_dq_decompress(stream):
    """g0 from a compressed stream held in a uint8 CUDA tensor
    (pmsz_codec_dq_decompress)."""
    import torch
    stream = stream.contiguous()
    p = Params()
    rc = lib().pmsz_codec_dq_stream_params(ctypes.c_void_p(stream.data_ptr()), stream.numel(),
                                           ctypes.byref(p))
    _check(rc, "pmsz_codec_dq_stream_params")
    dt = torch.float32 if p.dtype == 0 else torch.float64
    g0 = torch.empty((p.nz, p.ny, p.nx), dtype=dt, device=stream.device)
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_codec_dq_decompress(ctypes.c_void_p(stream.data_ptr()), stream.numel(),
                                        ctypes.c_void_p(g0.data_ptr()))
    _check(rc, "pmsz_codec_dq_decompress")
    return g0


def time_first_pass(f, g, abs_eb, iters=5, **kw):
    """(labeling seconds, detection-1 seconds), averaged, on CUDA tensors f, g
    (g is not modified)."""
    import torch
    p = make_params(f.shape, f.dtype, abs_eb, **kw)
    out = (ctypes.c_double * 2)()
    torch.cuda.current_stream().synchronize()
    rc = lib().pmsz_time_first_pass(ctypes.byref(p), ctypes.c_void_p(f.data_ptr()),
                                    ctypes.c_void_p(g.data_ptr()), int(iters), out)
    _check(rc, "pmsz_time_first_pass")
    return out[0], out[1]


def test_step_budget(f, g, d, abs_eb, q, clamp_mode=CLAMP_BOUND_SAFE):
    """Test hook: the step budgets (pmsz_test_step_budget) of CUDA tensors f, g
    (float32 or float64) and uint8 counters d; returns a uint8 tensor."""
    import torch
    out = torch.empty_like(d)
    rc = lib().pmsz_test_step_budget(_dtype_code(f.dtype), f.numel(),
                                     ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                                     ctypes.c_void_p(d.data_ptr()), float(abs_eb), int(q),
                                     int(clamp_mode), ctypes.c_void_p(out.data_ptr()))
    _check(rc, "pmsz_test_step_budget")
    return out


def version():
    return lib().pmsz_version().decode()
