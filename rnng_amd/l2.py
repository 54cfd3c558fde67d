"""Host-side interface of the L2 epsilon-graph path.

Mirrors what the reference's drivers do with the search API
(include/search.hpp:223-264, :266-340, :389-407) for one GPU's point set:
all pairs within eps of each other, weights equal to the reference's float
L2 distance.  Arrays are numpy; the compute runs in librnng_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MODE_GRID, MODE_MFMA, MODE_PLAIN, Stats, check


def _f32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("points must be a 2-D array [n, d]")
    return a


def _mode(mode) -> int:
    if isinstance(mode, str):
        return {"plain": MODE_PLAIN, "mfma": MODE_MFMA, "grid": MODE_GRID}[mode]
    return int(mode)


def sq_threshold(eps: float) -> np.float32:
    """Largest float T with sqrtf(T) <= eps."""
    return np.float32(_lib.lib().rnng_sq_threshold(np.float32(eps)))


class Dataset:
    """A point set resident in HBM (rnng_dataset_t)."""

    def __init__(self, points):
        pts = _f32(points)
        self.n, self.d = pts.shape
        h = C.c_void_p()
        check(_lib.lib().rnng_dataset_create_f32(pts.ctypes.data_as(C.c_void_p), self.n, self.d,
                                                 C.byref(h)))
        self._h = h

    @classmethod
    def from_handle(cls, h) -> "Dataset":
        """Adopt a dataset handle created by the library (e.g. from a file)."""
        self = cls.__new__(cls)
        n, d = C.c_int64(), C.c_int32()
        check(_lib.lib().rnng_dataset_shape(h, C.byref(n), C.byref(d)))
        self.n, self.d = n.value, d.value
        self._h = h
        return self

    def to_numpy(self) -> np.ndarray:
        """The points copied back from HBM."""
        out = np.empty((self.n, self.d), np.float32)
        check(_lib.lib().rnng_dataset_copy_f32(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            _lib.lib().rnng_dataset_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graph(self, eps: float, mode="mfma", stats: Stats | None = None) -> "Graph":
        h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_graph_build(self._h, np.float32(eps), _mode(mode), C.byref(h),
                                          C.byref(st)))
        return Graph(h, st)

    def bipartite(self, other: "Dataset", eps: float, mode="mfma",
                  stats: Stats | None = None) -> "Graph":
        h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_graph_build_bipartite(self._h, other._h, np.float32(eps), _mode(mode),
                                                    C.byref(h), C.byref(st)))
        return Graph(h, st)


class Graph:
    """Device-resident edge list (rnng_graph_t)."""

    def __init__(self, h, stats: Stats):
        self._h = h
        self.stats = stats

    @property
    def num_edges(self) -> int:
        n = C.c_int64()
        check(_lib.lib().rnng_graph_num_edges(self._h, C.byref(n)))
        return n.value

    def edges(self, sorted: bool = True):
        """(ij int64 [E,2], w float32 [E]), sorted by (i, j) if asked."""
        ne = self.num_edges
        ij = np.empty((ne, 2), np.int64)
        w = np.empty(ne, np.float32)
        check(_lib.lib().rnng_graph_copy(self._h, ij.ctypes.data_as(C.c_void_p),
                                         w.ctypes.data_as(C.c_void_p), ne, int(sorted)))
        return ij, w

    def csr(self, n_vertices: int, self_loops: bool = False):
        """Symmetric CSR built on the device: (rowptr int64 [n+1], col int32,
        w float32), columns ascending within each row."""
        L = _lib.lib()
        need = C.c_int64()
        check(L.rnng_graph_copy_csr(self._h, int(n_vertices), int(self_loops), None, None, None, 0,
                                    C.byref(need)))
        rowptr = np.empty(int(n_vertices) + 1, np.int64)
        col = np.empty(need.value, np.int32)
        w = np.empty(need.value, np.float32)
        check(L.rnng_graph_copy_csr(self._h, int(n_vertices), int(self_loops), rowptr.ctypes.data_as(C.c_void_p),
                                    col.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p), need.value,
                                    C.byref(need)))
        return rowptr, col, w

    def close(self) -> None:
        if self._h:
            _lib.lib().rnng_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def eps_pairs(points, eps: float, mode="mfma", stats: Stats | None = None):
    """Self join: all i < j with L2 <= eps, sorted by (i, j).  Returns (ij, w)."""
    ds = Dataset(points)
    try:
        g = ds.graph(eps, mode, stats)
        try:
            return g.edges(sorted=True)
        finally:
            g.close()
    finally:
        ds.close()


def eps_bipartite(a, b, eps: float, mode="mfma", stats: Stats | None = None):
    """All (i, j), i in a, j in b, with L2 <= eps, sorted by (i, j)."""
    A, B = Dataset(a), Dataset(b)
    try:
        if A.d != B.d:
            raise ValueError("dimension mismatch")
        g = A.bipartite(B, eps, mode, stats)
        try:
            return g.edges(sorted=True)
        finally:
            g.close()
    finally:
        A.close()
        B.close()


def eps_pairs_oneshot(points, eps: float, mode="mfma"):
    """The count-then-fill host entry point (rnng_l2_eps_pairs_f32): first call
    with cap=0 learns the size, second call fills."""
    pts = _f32(points)
    n, d = pts.shape
    L = _lib.lib()
    needed = C.c_int64()
    rc = L.rnng_l2_eps_pairs_f32(pts.ctypes.data_as(C.c_void_p), n, d, np.float32(eps), _mode(mode),
                                 None, None, 0, C.byref(needed), None)
    if rc not in (_lib.RNNG_OK, _lib.RNNG_ERR_CAPACITY):
        check(rc)
    cap = needed.value
    ij = np.empty((cap, 2), np.int64)
    w = np.empty(cap, np.float32)
    if cap:
        check(L.rnng_l2_eps_pairs_f32(pts.ctypes.data_as(C.c_void_p), n, d, np.float32(eps),
                                      _mode(mode), ij.ctypes.data_as(C.c_void_p),
                                      w.ctypes.data_as(C.c_void_p), cap, C.byref(needed), None))
    return ij, w


def mfma_layout():
    """(rows, cols) int32 [64,16]: output position of each (lane, register)."""
    rows = np.empty((64, 16), np.int32)
    cols = np.empty((64, 16), np.int32)
    check(_lib.lib().rnng_debug_mfma_layout(rows.ctypes.data_as(C.c_void_p),
                                            cols.ctypes.data_as(C.c_void_p)))
    return rows, cols


def device_sqrt(x) -> np.ndarray:
    """The device's float square root (diagnostic)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    check(_lib.lib().rnng_debug_sqrt(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                     x.size))
    return out
