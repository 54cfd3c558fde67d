"""Point files of the reference's formats (dataset_io.py of gabe-raulet/RNNG):

* ``.u8bin``  uint32 n, uint32 d, then n*d uint8   (dataset_io.py:48-53)
* ``.fbin``   uint32 n, uint32 d, then n*d float32 (dataset_io.py:55-60)
* ``.fvecs``  n records of int32 d + d float32     (dataset_io.py:62-70)

``read_file``/``info_file``/``write_file`` dispatch on the extension as the
reference's do (ValueError for any other); ``start`` and ``count`` select
points [start, start + count).  ``load_device`` reads a file straight into
HBM through the library (pinned, chunked, asynchronous upload; u8 values
widened to float32 on the device) and returns an ``l2.Dataset``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check


def info_u8bin(fname: str):
    n, d = np.fromfile(fname, count=2, dtype=np.uint32)
    return n, d, os.path.getsize(fname), "uint8"


def info_fbin(fname: str):
    n, d = np.fromfile(fname, count=2, dtype=np.uint32)
    return n, d, os.path.getsize(fname), "float32"


def info_fvecs(fname: str):
    size = os.path.getsize(fname)
    d = np.fromfile(fname, dtype=np.int32, count=1)[0]
    if size % (4 * (d + 1)) != 0:
        raise ValueError(f"malformed fvecs file '{fname}'")
    return size // (4 * (d + 1)), d, size, "float32"


def _bin(fname: str, dtype, start: int, count: int | None) -> np.ndarray:
    with open(fname, "rb") as f:
        n, d = np.fromfile(f, count=2, dtype=np.uint32)
        n = int(n) - start if count is None else count
        item = np.dtype(dtype).itemsize
        arr = np.fromfile(f, count=n * int(d), dtype=dtype, offset=start * int(d) * item)
    return arr.reshape(n, int(d))


def read_u8bin(fname: str, start: int = 0, count: int | None = None) -> np.ndarray:
    return _bin(fname, np.uint8, start, count)


def read_fbin(fname: str, start: int = 0, count: int | None = None) -> np.ndarray:
    return _bin(fname, np.float32, start, count)


def read_fvecs(fname: str, start: int = 0, count: int | None = None) -> np.ndarray:
    n, d, _, _ = info_fvecs(fname)
    n = n - start if count is None else count
    arr = np.fromfile(fname, count=n * (d + 1), dtype=np.float32, offset=start * 4 * (d + 1))
    return arr.reshape(-1, d + 1)[:, 1:].copy()


def write_u8bin(fname: str, points: np.ndarray) -> None:
    with open(fname, "wb") as f:
        np.array(points.shape, dtype=np.uint32).tofile(f)
        np.ascontiguousarray(points, dtype=np.uint8).tofile(f)


def write_fbin(fname: str, points: np.ndarray) -> None:
    with open(fname, "wb") as f:
        np.array(points.shape, dtype=np.uint32).tofile(f)
        np.ascontiguousarray(points, dtype=np.float32).tofile(f)


def write_fvecs(fname: str, points: np.ndarray) -> None:
    pts = np.ascontiguousarray(points, dtype=np.float32)
    rec = np.empty((pts.shape[0], pts.shape[1] + 1), np.float32)
    rec[:, 0] = np.int32(pts.shape[1]).view(np.float32)
    rec[:, 1:] = pts
    rec.tofile(fname)


_EXT = {".u8bin": (info_u8bin, read_u8bin, write_u8bin),
        ".fbin": (info_fbin, read_fbin, write_fbin),
        ".fvecs": (info_fvecs, read_fvecs, write_fvecs)}


def _kind(fname: str):
    for ext, fns in _EXT.items():
        if fname.endswith(ext):
            return fns
    raise ValueError(f"cannot read file '{fname}': unknown extension")


def info_file(fname: str):
    """(number of points, dimension, file size in bytes, atom type)."""
    return _kind(fname)[0](fname)


def read_file(fname: str, start: int = 0, count: int | None = None) -> np.ndarray:
    return _kind(fname)[1](fname, start, count)


def write_file(fname: str, points: np.ndarray) -> None:
    _kind(fname)[2](fname, points)


def load_device(fname: str, start: int = 0, count: int | None = None):
    """The points of a file, read straight into HBM (u8 values as float32)."""
    from .l2 import Dataset
    h = C.c_void_p()
    check(_lib.lib().rnng_dataset_read_file(os.fsencode(fname), int(start), -1 if count is None else int(count),
                                            C.byref(h)))
    return Dataset.from_handle(h)


def device_info(fname: str):
    """(points, dimension, kind) of a file as the library reads its header
    (kind 0: float32, 1: uint8)."""
    n, d, k = C.c_int64(), C.c_int32(), C.c_int32()
    check(_lib.lib().rnng_file_info(os.fsencode(fname), C.byref(n), C.byref(d), C.byref(k)))
    return n.value, d.value, k.value
