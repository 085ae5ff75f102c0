"""pmsz_anonymous_amd -- MI355X-native extrema + Morse-Smale segmentation
correction (the capability of pMSz), behind a C-ABI (include/pmsz.h).

Modules:
  engine   ctypes view of libpmsz.so: correct() (lazy, sync or naive-MSz mode),
           correct_batch(), codes(), labels(), edit_symbols(), encode_edits() (the
           edit side channel), codec_dq() (the GPU compressor front end)
  fieldio  ctypes view of libpmsz_io.so: Perlin fields, SZ-like / ZFP codecs, the
           zstd edit files (write_edits / read_edits), raw I/O
  serve    serving loop over contexts
  build    in-tree build recipe of the native libraries
"""
from . import engine, fieldio  # noqa: F401
from .engine import (PmszError, codec_dq, codes, correct, correct_batch,  # noqa: F401
                     edit_symbols, encode_edits, labels)

__all__ = ["engine", "fieldio", "PmszError", "correct", "correct_batch", "codes", "labels",
           "edit_symbols", "encode_edits", "codec_dq"]
