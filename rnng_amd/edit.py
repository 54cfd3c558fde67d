"""Host-side interface of the edit-distance epsilon-graph path.

Strings are a CSR (atoms uint8, offsets int64[n+1]) -- the layout the
reference's PointContainer<char> holds after read_seqs
(include/point.hpp:170-205).  Edges are the pairs i < j whose Levenshtein
distance (drivers/systolic.cpp:271-299) is <= eps, weight = float(distance).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Stats, check
from .l2 import Graph


def as_csr(strings):
    """list of bytes -> (atoms uint8, offsets int64)."""
    offs = np.zeros(len(strings) + 1, np.int64)
    offs[1:] = np.cumsum([len(s) for s in strings])
    atoms = np.frombuffer(b"".join(strings), np.uint8) if strings else np.zeros(0, np.uint8)
    return atoms, offs


class StringSet:
    """A string set resident in HBM (rnng_strings_t)."""

    def __init__(self, atoms, offsets):
        self.atoms = np.ascontiguousarray(atoms, dtype=np.uint8)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.n = self.offsets.shape[0] - 1
        h = C.c_void_p()
        ap = self.atoms.ctypes.data_as(C.c_void_p) if self.atoms.size else None
        check(_lib.lib().rnng_strings_create(ap, self.offsets.ctypes.data_as(C.c_void_p), self.n,
                                             C.byref(h)))
        self._h = h

    @classmethod
    def from_file(cls, path: str, start: int = 0, count: int | None = None) -> "StringSet":
        """A .seqs file read straight into HBM (rnng_strings_read_file)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(_lib.lib().rnng_strings_read_file(path.encode(), int(start), -1 if count is None else int(count),
                                                C.byref(h)))
        self._h = h
        n, na = C.c_int64(), C.c_int64()
        check(_lib.lib().rnng_strings_size(h, C.byref(n), C.byref(na)))
        self.n = n.value
        self.atoms = np.empty(na.value, np.uint8)
        self.offsets = np.empty(self.n + 1, np.int64)
        check(_lib.lib().rnng_strings_copy(h, self.atoms.ctypes.data_as(C.c_void_p),
                                           self.offsets.ctypes.data_as(C.c_void_p)))
        return self

    def close(self) -> None:
        if self._h:
            _lib.lib().rnng_strings_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graph(self, eps: float, stats: Stats | None = None) -> Graph:
        h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_edit_graph_build(self._h, np.float32(eps), C.byref(h), C.byref(st)))
        return Graph(h, st)

    def bipartite(self, other: "StringSet", eps: float, stats: Stats | None = None) -> Graph:
        """All (i in self, j in other) with edit distance <= eps."""
        h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_edit_graph_build_bipartite(self._h, other._h, np.float32(eps), C.byref(h),
                                                         C.byref(st)))
        return Graph(h, st)


def eps_pairs(atoms, offsets, eps: float, stats: Stats | None = None):
    """All i < j with edit distance <= eps, sorted.  Returns (ij int64 [E,2],
    dist int32 [E])."""
    S = StringSet(atoms, offsets)
    try:
        g = S.graph(eps, stats)
        try:
            ij, w = g.edges(sorted=True)
        finally:
            g.close()
    finally:
        S.close()
    return ij, w.astype(np.int32)
