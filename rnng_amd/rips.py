"""Host-side interface of the Rips filtration path.

rips_filtration(points, eps, maxdim) returns the merged filtration that the
reference's rips driver writes (include/point.hpp:1134-1267 per cell,
include/simplex.hpp:130-196 to merge): (value float32 [S], id int64 [S]),
sorted by (value, id).  Simplex ids follow include/simplex.hpp:5-19.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Stats, check
from .l2 import Dataset


def rips_filtration(points, eps: float, maxdim: int = 2, stats: Stats | None = None):
    ds = points if isinstance(points, Dataset) else Dataset(points)
    st = stats if stats is not None else Stats()
    h = C.c_void_p()
    L = _lib.lib()
    try:
        check(L.rnng_rips_build(ds.handle, np.float32(eps), int(maxdim), C.byref(h), C.byref(st)))
        try:
            n = C.c_int64()
            per = np.zeros(4, np.int64)
            check(L.rnng_filtration_size(h, C.byref(n), per.ctypes.data_as(C.c_void_p)))
            val = np.empty(n.value, np.float32)
            ids = np.empty(n.value, np.int64)
            check(L.rnng_filtration_copy(h, val.ctypes.data_as(C.c_void_p),
                                         ids.ctypes.data_as(C.c_void_p), n.value))
        finally:
            L.rnng_filtration_destroy(h)
    finally:
        if ds is not points:
            ds.close()
    return val, ids


def simplex_dims(ids: np.ndarray) -> np.ndarray:
    return (ids.astype(np.int64) >> 60) & 0xF


def write_filtration(fname: str, value: np.ndarray, ids: np.ndarray) -> None:
    """The merged filtration text file of include/simplex.hpp:175-192
    (use_ids form): one "%f\\t%lld" line per simplex."""
    with open(fname, "w") as f:
        for v, i in zip(value.tolist(), ids.tolist()):
            f.write("%f\t%d\n" % (v, i))
