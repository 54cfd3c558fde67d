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


class Filtration:
    """Device-resident filtration (rnng_filtration_t) of `ds` at eps."""

    def __init__(self, ds: Dataset, eps: float, maxdim: int = 2, stats: Stats | None = None):
        self._h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_rips_build(ds.handle, np.float32(eps), int(maxdim), C.byref(self._h),
                                         C.byref(st)))

    def size(self):
        """(total simplices, per-dimension counts [4])."""
        n = C.c_int64()
        per = np.zeros(4, np.int64)
        check(_lib.lib().rnng_filtration_size(self._h, C.byref(n), per.ctypes.data_as(C.c_void_p)))
        return n.value, per

    def copy(self):
        """(value float32 [S], id int64 [S]) in filtration order."""
        n, _ = self.size()
        val = np.empty(n, np.float32)
        ids = np.empty(n, np.int64)
        check(_lib.lib().rnng_filtration_copy(self._h, val.ctypes.data_as(C.c_void_p),
                                              ids.ctypes.data_as(C.c_void_p), n))
        return val, ids

    def close(self) -> None:
        if self._h:
            _lib.lib().rnng_filtration_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def rips_filtration(points, eps: float, maxdim: int = 2, stats: Stats | None = None):
    ds = points if isinstance(points, Dataset) else Dataset(points)
    try:
        with Filtration(ds, eps, maxdim, stats) as f:
            return f.copy()
    finally:
        if ds is not points:
            ds.close()


def simplex_dims(ids: np.ndarray) -> np.ndarray:
    return (ids.astype(np.int64) >> 60) & 0xF


def write_filtration(fname: str, value: np.ndarray, ids: np.ndarray) -> None:
    """The merged filtration text file of include/simplex.hpp:175-192
    (use_ids form): one "%f\\t%lld" line per simplex."""
    with open(fname, "w") as f:
        for v, i in zip(value.tolist(), ids.tolist()):
            f.write("%f\t%d\n" % (v, i))
