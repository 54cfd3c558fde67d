"""Host-side interface of the Voronoi landmark path.

Mirrors the landmark driver's use of the point API (drivers/landmark.cpp:
110-210): landmarks (selection_sample, include/utils.hpp:1-29, or
farthest-point sampling), the Voronoi diagram (include/point.hpp:535-559),
coalesced cells (:561-758), ghost points (:760-929) and the per-cell epsilon
query, on one GPU.  The compute runs in librnng_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MODE_MFMA, Stats, check
from .l2 import Dataset, Graph, _f32, _mode


def _vp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def selection_sample(range_: int, size: int, seed: int) -> np.ndarray:
    """Knuth's selection sampling with the reference's engine: `size` sorted
    ids from [0, range_)."""
    out = np.empty(max(size, 0), np.int64)
    check(_lib.lib().rnng_selection_sample(int(range_), int(size), _vp(out), int(seed)))
    return out


def voronoi_assign(points, centers):
    """Nearest center of each point (ties to the later center) and the distance."""
    pts, ctr = _f32(points), _f32(centers)
    if pts.shape[1] != ctr.shape[1]:
        raise ValueError("dimension mismatch")
    n, m = pts.shape[0], ctr.shape[0]
    cell = np.empty(n, np.int64)
    dist = np.empty(n, np.float32)
    check(_lib.lib().rnng_voronoi_assign_f32(_vp(pts), n, _vp(ctr), m, pts.shape[1], _vp(cell),
                                             _vp(dist)))
    return cell, dist


class Voronoi:
    """Voronoi cells of a device-resident dataset (rnng_voronoi_t).

    Voronoi(ds, centers=ids) uses the given center point ids (cell c is
    centers[c]); Voronoi(ds, fps=m, first=i) picks m farthest-point
    landmarks starting at point i.  The dataset must stay open while the
    handle lives."""

    def __init__(self, ds: Dataset, centers=None, *, fps: int | None = None, first: int = 0):
        h = C.c_void_p()
        L = _lib.lib()
        if centers is not None:
            c = np.ascontiguousarray(centers, dtype=np.int64)
            check(L.rnng_voronoi_create(ds.handle, _vp(c), c.shape[0], C.byref(h)))
            self.m = int(c.shape[0])
        elif fps is not None:
            check(L.rnng_voronoi_create_fps(ds.handle, int(fps), int(first), C.byref(h)))
            self.m = int(fps)
        else:
            raise ValueError("give centers or fps")
        self._h = h
        self._ds = ds
        self.n = ds.n

    def centers(self) -> np.ndarray:
        out = np.empty(self.m, np.int64)
        check(_lib.lib().rnng_voronoi_centers(self._h, _vp(out)))
        return out

    def assignment(self):
        cell = np.empty(self.n, np.int64)
        dist = np.empty(self.n, np.float32)
        check(_lib.lib().rnng_voronoi_assignment(self._h, _vp(cell), _vp(dist)))
        return cell, dist

    def cells(self):
        ptr = np.empty(self.m + 1, np.int64)
        ids = np.empty(self.n, np.int64)
        check(_lib.lib().rnng_voronoi_cells(self._h, _vp(ptr), _vp(ids)))
        return ptr, ids

    def graph(self, eps: float, mode="mfma", stats: Stats | None = None) -> Graph:
        h = C.c_void_p()
        st = stats if stats is not None else Stats()
        check(_lib.lib().rnng_voronoi_graph_build(self._h, np.float32(eps), _mode(mode), C.byref(h),
                                                  C.byref(st)))
        return Graph(h, st)

    def set_options(self, use_graph: int = -1, autotune: int = -1) -> None:
        """1 on, 0 off, -1 default (on) for the captured fill pass and the
        filter autotuner of the next builds."""
        check(_lib.lib().rnng_voronoi_set_options(self._h, int(use_graph), int(autotune)))

    def tuning(self):
        """(variant, xcd_order, timings ms of the 8 tuned configurations)."""
        var, xcd = C.c_int32(), C.c_int32()
        ms = np.zeros(8, np.float64)
        check(_lib.lib().rnng_voronoi_tuning(self._h, C.byref(var), C.byref(xcd), _vp(ms)))
        return var.value, xcd.value, ms

    def set_schedule(self, order: int = -1, sched: int = -1) -> None:
        """Force the row order (0 cells in landmark order, 1 cells along the
        center chain, 2 flat: points in their own order, every pair; runs only)
        and work form (0 runs, 1 tile lists, 2 tile lists on a persistent
        queue); -1: tuner."""
        check(_lib.lib().rnng_voronoi_set_schedule(self._h, int(order), int(sched)))

    def schedule(self):
        """(order, form, [(order, form, xcd, ms, shape) of every tuned
        configuration]); shape: the BF16 variant or the int8 code (0: the cell
        default)."""
        o, s, nc = C.c_int32(), C.c_int32(), C.c_int32()
        cfg = np.zeros(64, np.int32)
        ms = np.zeros(64, np.float64)
        check(_lib.lib().rnng_voronoi_schedule(self._h, C.byref(o), C.byref(s), _vp(cfg), _vp(ms), 64,
                                               C.byref(nc)))
        k = min(nc.value, 64)
        return o.value, s.value, [((int(c) >> 8) & 255, (int(c) >> 4) & 15, int(c) & 1, float(t), int(c) >> 16)
                                  for c, t in zip(cfg[:k], ms[:k])]

    def ghosts(self) -> np.ndarray:
        """(cell, point) pairs of the last graph build, sorted."""
        n = C.c_int64()
        check(_lib.lib().rnng_voronoi_num_ghosts(self._h, C.byref(n)))
        out = np.empty((n.value, 2), np.int64)
        check(_lib.lib().rnng_voronoi_ghosts(self._h, _vp(out), n.value))
        return out

    def close(self) -> None:
        if self._h:
            _lib.lib().rnng_voronoi_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["selection_sample", "voronoi_assign", "Voronoi", "MODE_MFMA"]
