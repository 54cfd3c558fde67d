"""CPU oracle for the epsilon-neighbor-graph path (ctypes wrapper of liborc.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker.  The product (``rnng_amd``) never imports it.

Each function restates a reference routine; the C source (oracle.c / oracle.h)
cites the reference file:line for each and lists the checks that pin it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


def build() -> None:
    """Compile liborc.so with the pinned FP flags (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    lib.orc_l2sq.argtypes = [f32p, f32p, C.c_int64]
    lib.orc_l2sq.restype = C.c_float
    lib.orc_l2.argtypes = [f32p, f32p, C.c_int64]
    lib.orc_l2.restype = C.c_float
    lib.orc_sq_threshold.argtypes = [C.c_float]
    lib.orc_sq_threshold.restype = C.c_float
    lib.orc_l2_eps_pairs.argtypes = [f32p, C.c_int64, f32p, C.c_int64, C.c_int64, C.c_float,
                                     C.c_int, C.c_int, vp, vp, C.c_int64]
    lib.orc_l2_eps_pairs.restype = C.c_int64
    lib.orc_edit_distance.argtypes = [C.c_char_p, C.c_int64, C.c_char_p, C.c_int64]
    lib.orc_edit_distance.restype = C.c_int64
    lib.orc_edit_eps_pairs.argtypes = [vp, i64p, C.c_int64, C.c_float, C.c_int, vp, vp, C.c_int64]
    lib.orc_edit_eps_pairs.restype = C.c_int64
    lib.orc_voronoi_assign.argtypes = [f32p, C.c_int64, f32p, C.c_int64, C.c_int64, C.c_int,
                                       i64p, f32p]
    lib.orc_voronoi_assign.restype = None
    lib.orc_ghosts.argtypes = [f32p, C.c_int64, C.c_int64, i64p, f32p, f32p, C.c_int64,
                               C.c_float, C.c_int, vp, C.c_int64]
    lib.orc_ghosts.restype = C.c_int64
    lib.orc_fps.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int, i64p, i64p,
                            f32p]
    lib.orc_fps.restype = None
    lib.orc_binom.argtypes = [C.c_int64, C.c_int64]
    lib.orc_binom.restype = C.c_int64
    lib.orc_simplex_id.argtypes = [i64p, C.c_int64]
    lib.orc_simplex_id.restype = C.c_int64
    lib.orc_rips_filtration.argtypes = [f32p, C.c_int64, C.c_int64, C.c_float, C.c_int64,
                                        C.c_int, vp, vp, C.c_int64]
    lib.orc_rips_filtration.restype = C.c_int64
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def l2sq(p, q) -> np.float32:
    p, q = _f32(p), _f32(q)
    return np.float32(_load().orc_l2sq(p, q, p.shape[-1]))


def l2(p, q) -> np.float32:
    p, q = _f32(p), _f32(q)
    return np.float32(_load().orc_l2(p, q, p.shape[-1]))


def sq_threshold(eps: float) -> np.float32:
    return np.float32(_load().orc_sq_threshold(np.float32(eps)))


def l2_eps_pairs(X, eps, Y=None, nthreads: int = 0, cap: int = 0):
    """Brute-force epsilon pairs.  Self join (Y is None): i < j.  Returns
    (ij int64 [E,2], w float32 [E]) in row-major order.  cap > 0 is a size
    guess: when the count fits, the pairs are computed in one pass instead of
    a counting pass and a filling pass."""
    lib = _load()
    X = _f32(X)
    self_join = Y is None
    Y = X if self_join else _f32(Y)
    nx, d = X.shape
    ny = Y.shape[0]
    eps = np.float32(eps)
    if cap > 0:
        ij = np.empty((cap, 2), np.int64)
        w = np.empty(cap, np.float32)
        n = lib.orc_l2_eps_pairs(X, nx, Y, ny, d, eps, int(self_join), nthreads, _ptr(ij), _ptr(w), cap)
        if n <= cap:
            return ij[:n].copy(), w[:n].copy()
    else:
        n = lib.orc_l2_eps_pairs(X, nx, Y, ny, d, eps, int(self_join), nthreads, None, None, 0)
    ij = np.empty((n, 2), np.int64)
    w = np.empty(n, np.float32)
    lib.orc_l2_eps_pairs(X, nx, Y, ny, d, eps, int(self_join), nthreads, _ptr(ij), _ptr(w), n)
    return ij, w


def l2_eps_pairs_count(X, eps, Y=None, nthreads: int = 0) -> int:
    lib = _load()
    X = _f32(X)
    self_join = Y is None
    Y = X if self_join else _f32(Y)
    return int(lib.orc_l2_eps_pairs(X, X.shape[0], Y, Y.shape[0], X.shape[1], np.float32(eps),
                                    int(self_join), nthreads, None, None, 0))


def edit_distance(s: bytes, t: bytes) -> int:
    return int(_load().orc_edit_distance(s, len(s), t, len(t)))


def edit_eps_pairs(atoms: np.ndarray, offsets: np.ndarray, eps, nthreads: int = 0, cap: int = 0):
    """Brute-force edit-distance pairs i < j with dist <= eps.  Returns
    (ij int64 [E,2], dist int32 [E]).  cap > 0: one pass when the count fits."""
    lib = _load()
    atoms = np.ascontiguousarray(atoms, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    n = offsets.shape[0] - 1
    eps = np.float32(eps)
    if cap > 0:
        ij = np.empty((cap, 2), np.int64)
        dist = np.empty(cap, np.int32)
        cnt = lib.orc_edit_eps_pairs(_ptr(atoms), offsets, n, eps, nthreads, _ptr(ij), _ptr(dist), cap)
        if cnt <= cap:
            return ij[:cnt].copy(), dist[:cnt].copy()
    else:
        cnt = lib.orc_edit_eps_pairs(_ptr(atoms), offsets, n, eps, nthreads, None, None, 0)
    ij = np.empty((cnt, 2), np.int64)
    dist = np.empty(cnt, np.int32)
    lib.orc_edit_eps_pairs(_ptr(atoms), offsets, n, eps, nthreads, _ptr(ij), _ptr(dist), cnt)
    return ij, dist


def voronoi_assign(pts, centers, nthreads: int = 0):
    lib = _load()
    pts, centers = _f32(pts), _f32(centers)
    n, d = pts.shape
    cell = np.empty(n, np.int64)
    dist = np.empty(n, np.float32)
    lib.orc_voronoi_assign(pts, n, centers, centers.shape[0], d, nthreads, cell, dist)
    return cell, dist


def ghosts(pts, cell, delta, centers, eps, nthreads: int = 0):
    """Ghost pairs (cell, point) sorted by (cell, point)."""
    lib = _load()
    pts, centers, delta = _f32(pts), _f32(centers), _f32(delta)
    cell = np.ascontiguousarray(cell, dtype=np.int64)
    n, d = pts.shape
    m = centers.shape[0]
    cnt = lib.orc_ghosts(pts, n, d, cell, delta, centers, m, np.float32(eps), nthreads, None, 0)
    out = np.empty((cnt, 2), np.int64)
    lib.orc_ghosts(pts, n, d, cell, delta, centers, m, np.float32(eps), nthreads, _ptr(out), cnt)
    return out


def fps(pts, m: int, first: int = 0, nthreads: int = 0):
    """Farthest-point landmarks: (centers int64[m], cell int64[n], dist float32[n])."""
    lib = _load()
    pts = _f32(pts)
    n, d = pts.shape
    centers = np.empty(m, np.int64)
    cell = np.empty(n, np.int64)
    dist = np.empty(n, np.float32)
    lib.orc_fps(pts, n, d, m, first, nthreads, centers, cell, dist)
    return centers, cell, dist


def voronoi_cells(cell, centers, m: int):
    """Cells in the reference's coalesce order for one rank
    (include/point.hpp:721-754): members in ascending id, then the center
    swapped with the first member.  Returns (ptr int64[m+1], ids int64[n])."""
    cell = np.asarray(cell, dtype=np.int64)
    order = np.lexsort((np.arange(cell.shape[0]), cell))
    ptr = np.zeros(m + 1, np.int64)
    ptr[1:] = np.cumsum(np.bincount(cell, minlength=m))
    ids = order.astype(np.int64)
    for c in range(m):
        lo, hi = ptr[c], ptr[c + 1]
        seg = ids[lo:hi]
        hit = np.nonzero(seg == centers[c])[0]
        if hit.size:
            k = hit[0]
            seg[0], seg[k] = seg[k], seg[0]
    return ptr, ids


_MINSTD_M = 2147483647          # minstd_rand0 modulus 2^31 - 1
_MINSTD_A = 16807               # minstd_rand0 multiplier
_MINSTD_R = float(_MINSTD_M - 1)  # engine range max - min + 1 (min 1, max 2^31 - 2)


def selection_sample(range_: int, size: int, seed: int):
    """Knuth's Algorithm S as the reference runs it (include/utils.hpp:1-29),
    restated from the published definitions of its library pieces:
    std::default_random_engine = minstd_rand0 (x <- 16807 x mod 2^31-1; a
    seed s becomes s mod (2^31-1), 0 becoming 1), generate_canonical<double,
    53> over a 2^31-2 range = two draws g1, g2 combined as
    ((g1-1) + (g2-1) R) / R^2 with R = 2^31-2, all in double, clamped below 1,
    and uniform_real_distribution(0, b) = canonical * b with
    b = nextafter(1, +inf).  Python floats are IEEE doubles, so every
    rounding step is the library's."""
    import math
    x = seed % _MINSTD_M
    if x == 0:
        x = 1
    b = math.nextafter(1.0, math.inf)
    rr = _MINSTD_R * _MINSTD_R
    out = []
    pos = 0
    for i in range(range_):
        if pos == size:
            break
        x = (_MINSTD_A * x) % _MINSTD_M
        g1 = x
        x = (_MINSTD_A * x) % _MINSTD_M
        g2 = x
        s = float(g1 - 1) * 1.0
        s = s + float(g2 - 1) * _MINSTD_R
        c = s / rr
        if c >= 1.0:
            c = math.nextafter(1.0, 0.0)
        u = c * (b - 0.0) + 0.0
        if float(range_ - i) * u < float(size - pos):
            out.append(i)
            pos += 1
    return np.asarray(out, dtype=np.int64)


def binom(n: int, k: int) -> int:
    return int(_load().orc_binom(n, k))


def simplex_id(verts) -> int:
    v = np.ascontiguousarray(verts, dtype=np.int64)
    return int(_load().orc_simplex_id(v, v.shape[0]))


def rips_filtration(pts, eps, maxdim: int = 2, nthreads: int = 0, cap: int = 0):
    """Sorted (value float32 [S], id int64 [S]) of the Rips filtration.
    cap > 0: one pass when the count fits."""
    lib = _load()
    pts = _f32(pts)
    n, d = pts.shape
    if cap > 0:
        val = np.empty(cap, np.float32)
        ids = np.empty(cap, np.int64)
        cnt = lib.orc_rips_filtration(pts, n, d, np.float32(eps), maxdim, nthreads, _ptr(val), _ptr(ids), cap)
        if cnt <= cap:
            return val[:cnt].copy(), ids[:cnt].copy()
    else:
        cnt = lib.orc_rips_filtration(pts, n, d, np.float32(eps), maxdim, nthreads, None, None, 0)
    val = np.empty(cnt, np.float32)
    ids = np.empty(cnt, np.int64)
    lib.orc_rips_filtration(pts, n, d, np.float32(eps), maxdim, nthreads, _ptr(val), _ptr(ids), cnt)
    return val, ids


# ---------------------------------------------------------------------------
# Cost-model perceptron (no reference counterpart): float64 numpy restatement
# of the library's 4 -> 64 -> 64 -> 2 ReLU network and its mean-squared-error
# gradient.  Pinned by central finite differences (tests/test_costmodel_cpu.py).
# ---------------------------------------------------------------------------
MLP_I, MLP_H, MLP_O = 4, 64, 2
MLP_NPARAMS = MLP_H * MLP_I + MLP_H + MLP_H * MLP_H + MLP_H + MLP_O * MLP_H + MLP_O


def mlp_unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in [(MLP_H, MLP_I), (MLP_H,), (MLP_H, MLP_H), (MLP_H,), (MLP_O, MLP_H), (MLP_O,)]:
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    return out


def mlp_forward(p, x):
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


MLP_UNDER = 4.0   # loss weight of an output below its target


def mlp_loss_grad(p, x, t, under=MLP_UNDER):
    """(loss, grad) of mean(w (y - t)^2) over batch and outputs, w = under
    (default 4) where y < t, else 1; ReLU'(0) = 0."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    B = x.shape[0]
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    r = y - t
    wr = np.where(r < 0, under, 1.0)
    loss = float(np.mean(wr * r * r))
    dy = 2.0 * wr * r / (B * MLP_O)
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3) * (h2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
