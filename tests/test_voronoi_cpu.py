"""CPU checks of the Voronoi landmark path's host logic and of the oracle
pieces it is judged against: the library's selection sampling against the
oracle's restatement of libstdc++'s engine and distributions, the oracle's
farthest-point sampling against a numpy restatement, and the cell grouping
rule.  No GPU call."""
import math

import numpy as np
import pytest

F = np.float32


def _np_l2(a, b):
    """Reference-order FP32 L2 of rows a[i] and b[j] (sequential sum over dims)."""
    acc = np.zeros(np.broadcast_shapes(a.shape[:-1], b.shape[:-1]), F)
    for k in range(a.shape[-1]):
        t = (a[..., k] - b[..., k]).astype(F)
        acc = (acc + (t * t).astype(F)).astype(F)
    return np.sqrt(acc).astype(F)


@pytest.mark.parametrize("rng_,size,seed", [(100, 10, 1), (1000, 37, 0), (2_000_000, 4096, 7),
                                            (50, 50, 3), (10, 0, 5), (123457, 1000, 2147483647),
                                            (1, 1, 9), (300_000, 64, 42)])
def test_selection_sample_matches_restatement(orc, rng_, size, seed):
    from rnng_amd import voronoi
    got = voronoi.selection_sample(rng_, size, seed)
    ref = orc.selection_sample(rng_, size, seed)
    assert np.array_equal(got, ref)
    assert got.shape == (size,)
    assert np.all(np.diff(got) > 0) and (size == 0 or (got[0] >= 0 and got[-1] < rng_))


def test_selection_sample_engine_known_answers(orc):
    """minstd_rand0's published check value: the 10000th draw from the default
    seed 1 is 1043618065 (C++ standard [rand.predef]); the restatement's
    engine reproduces it."""
    x = 1
    for _ in range(10000):
        x = (16807 * x) % 2147483647
    assert x == 1043618065
    # all-selected and nothing-selected edge cases follow from the algorithm
    assert np.array_equal(orc.selection_sample(7, 7, 11), np.arange(7))
    assert orc.selection_sample(7, 0, 11).size == 0


def test_selection_sample_rejects_bad_sizes():
    from rnng_amd import voronoi
    from rnng_amd._lib import RnngError
    with pytest.raises(RnngError):
        voronoi.selection_sample(5, 6, 1)


def test_selection_sample_is_uniformish(orc):
    """Inclusion frequency of each id is size/range (Algorithm S is exact)."""
    hits = np.zeros(40, np.int64)
    for seed in range(400):
        hits[orc.selection_sample(40, 10, seed)] += 1
    assert abs(hits.mean() - 100.0) < 1e-9
    assert hits.min() > 60 and hits.max() < 140


def _np_fps(pts, m, first):
    n = pts.shape[0]
    D = _np_l2(pts[:, None, :], pts[None, :, :])
    centers = [first]
    dist = np.full(n, np.finfo(F).max, F)
    cell = np.zeros(n, np.int64)
    for k in range(m):
        dd = D[centers[k]]
        upd = dd <= dist
        dist = np.where(upd, dd, dist)
        cell = np.where(upd, k, cell)
        if k + 1 < m:
            centers.append(int(np.argmax(dist)))   # first maximal index
    return np.array(centers), cell, dist


@pytest.mark.parametrize("n,d,m,first", [(300, 4, 20, 0), (257, 16, 33, 100), (64, 3, 64, 5)])
def test_oracle_fps_matches_numpy(orc, n, d, m, first):
    rng = np.random.default_rng(n + d)
    pts = rng.integers(0, 6, size=(n, d)).astype(F)   # small integers: many ties
    c, cell, dist = orc.fps(pts, m, first)
    c2, cell2, dist2 = _np_fps(pts, m, first)
    assert np.array_equal(c, c2)
    assert np.array_equal(cell, cell2)
    assert dist.tobytes() == dist2.tobytes()
    # the by-product assignment is the Voronoi assignment of the chosen centers
    cell3, dist3 = orc.voronoi_assign(pts, pts[c])
    assert np.array_equal(cell, cell3)
    assert dist.tobytes() == dist3.tobytes()


def test_oracle_cells_rule(orc):
    cell = np.array([2, 0, 2, 1, 0, 2, 2], np.int64)
    centers = np.array([4, 3, 5], np.int64)
    ptr, ids = orc.voronoi_cells(cell, centers, 3)
    assert ptr.tolist() == [0, 2, 3, 7]
    assert ids.tolist() == [4, 1, 3, 5, 2, 0, 6]   # cell 2: [0,2,5,6] -> center 5 to front
    # a center that is not a member of its own cell (duplicate center) leaves the order alone
    ptr, ids = orc.voronoi_cells(np.array([1, 1, 0]), np.array([0, 1]), 2)
    assert ids.tolist() == [2, 1, 0]


@pytest.mark.parametrize("scale,d,m,seed", [(1.0, 8, 40, 1), (1e-3, 16, 25, 2), (3e4, 5, 60, 3), (1.0, 33, 12, 4)])
def test_cell_adjacency_bound_keeps_every_pair_within_eps(orc, scale, d, m, seed):
    """The cell adjacency of the graph build, restated (FP32 center distances
    summed in order, R = a cell's largest FP32 member distance, the bound
    (R_a + R_b + eps)(1 + 1e-3) + 1e-15 compared in double): every pair the
    oracle finds within eps lies in one cell or in two adjacent cells, and on
    clustered data the bound still prunes some cell pairs."""
    rng = np.random.default_rng(seed)
    mu = rng.uniform(0, 10, size=(6, d))
    X = (mu[rng.integers(0, 6, 3000)] + 0.6 * rng.standard_normal((3000, d))) * scale
    X = X.astype(F)
    from rnng_amd import synth
    eps = F(synth.radius_for_degree(X, 15.0))
    ij, _ = orc.l2_eps_pairs(X, eps)
    centers = rng.choice(X.shape[0], m, replace=False)
    cell, delta = orc.voronoi_assign(X, X[centers])
    R = np.zeros(m, np.float64)
    np.maximum.at(R, cell, delta.astype(np.float64))
    adj = _adjacency(X[centers], R, eps)
    a, b = cell[ij[:, 0]], cell[ij[:, 1]]
    assert ij.shape[0] > 1000
    assert np.all((a == b) | adj[a, b])
    assert not adj.all()   # the bound prunes on this data


def _adjacency(C, R, eps, wrong=False):
    """The graph build's cell adjacency restated: FP32 center distances summed
    in order, min(R_a + R_b + eps, 2 (min(R_a, R_b) + eps)) with a 1e-3
    relative and 1e-15 absolute margin, compared in double.  wrong: the
    triangle bound without eps (a planted mistake)."""
    D = _np_l2(C[:, None, :], C[None, :, :]).astype(np.float64)
    Ra, Rb = R[:, None], R[None, :]
    if wrong:
        return D <= Ra + Rb
    return D <= np.minimum(Ra + Rb + float(eps), 2.0 * (np.minimum(Ra, Rb) + float(eps))) * (1.0 + 1e-3) + 1e-15


def test_cell_adjacency_bound_at_its_limit(orc):
    """Two cells on a line whose closest cross pair is exactly eps apart, the
    centers R_a + R_b + eps apart: the pair needs the two cells adjacent, which
    the bound keeps, and the same bound without eps (a planted mistake) does
    not."""
    X = np.zeros((82, 4), F)
    X[:41, 0] = np.arange(-20, 21)
    X[41:, 0] = np.arange(29, 70)
    X[:, 1] = np.tile([0.0, 1.0], 41)
    eps = orc.l2(X[40], X[41])
    ij, _ = orc.l2_eps_pairs(X, eps)
    centers = np.array([20, 61])
    cell, delta = orc.voronoi_assign(X, X[centers])
    R = np.zeros(2, np.float64)
    np.maximum.at(R, cell, delta.astype(np.float64))
    a, b = cell[ij[:, 0]], cell[ij[:, 1]]
    assert np.any(a != b)
    assert np.all((a == b) | _adjacency(X[centers], R, eps)[a, b])
    assert not np.all((a == b) | _adjacency(X[centers], R, eps, wrong=True)[a, b])
