"""pmsz_anonymous_amd -- MI355X-native extrema + Morse-Smale segmentation
correction (the capability of pMSz), behind a C-ABI (include/pmsz.h).

Modules:
  engine   ctypes view of libpmsz.so: correct(), codes(), labels(), edit_symbols()
  fieldio  ctypes view of libpmsz_io.so: Perlin fields, SZ-like / ZFP codecs, raw I/O
  build    in-tree build recipe of the native libraries
"""
from . import engine, fieldio  # noqa: F401
from .engine import PmszError, correct, correct_batch, codes, labels, edit_symbols  # noqa: F401

__all__ = ["engine", "fieldio", "PmszError", "correct", "codes", "labels", "edit_symbols"]
