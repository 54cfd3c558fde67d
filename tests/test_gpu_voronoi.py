"""GPU parity of the Voronoi landmark path through the C ABI: exact
assignment (ties to the later center), farthest-point landmarks, cells in
coalesce order, ghosts, and the batched per-cell query whose edge set must
equal the all-pairs oracle's bit for bit (ids and float weights)."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu
F = np.float32


def _mods():
    from rnng_amd import l2, voronoi
    return l2, voronoi


def _same_edges(ij, w, ij_ref, w_ref):
    assert ij.shape == ij_ref.shape, (ij.shape, ij_ref.shape)
    assert np.array_equal(ij, ij_ref)
    assert w.tobytes() == w_ref.tobytes()


@pytest.mark.parametrize("n,m,d,kind", [(5000, 7, 16, "rand"), (3000, 200, 96, "rand"),
                                        (2000, 130, 3, "int"), (1500, 1, 5, "rand"),
                                        (999, 999, 8, "int"), (4000, 300, 130, "rand")])
def test_assign_matches_oracle(orc, n, m, d, kind):
    _, vor = _mods()
    rng = np.random.default_rng(n * 7 + m)
    if kind == "int":   # small integer grid: many equal distances, duplicate centers
        pts = rng.integers(0, 3, size=(n, d)).astype(F)
    else:
        pts = rng.standard_normal((n, d)).astype(F)
    centers = pts[rng.integers(0, n, size=m)]
    flush_print(f"assign n={n} m={m} d={d} {kind}")
    cell, dist = vor.voronoi_assign(pts, centers)
    cell_ref, dist_ref = orc.voronoi_assign(pts, centers)
    assert np.array_equal(cell, cell_ref)
    assert dist.tobytes() == dist_ref.tobytes()


@pytest.mark.parametrize("n,d,m,first", [(20000, 16, 64, 0), (5000, 96, 100, 17), (3000, 3, 300, 2999)])
def test_fps_matches_oracle(orc, n, d, m, first):
    l2, vor = _mods()
    rng = np.random.default_rng(d)
    pts = rng.standard_normal((n, d)).astype(F)
    flush_print(f"fps n={n} d={d} m={m}")
    ds = l2.Dataset(pts)
    v = vor.Voronoi(ds, fps=m, first=first)
    c_ref, cell_ref, dist_ref = orc.fps(pts, m, first)
    assert np.array_equal(v.centers(), c_ref)
    cell, dist = v.assignment()
    assert np.array_equal(cell, cell_ref)
    assert dist.tobytes() == dist_ref.tobytes()
    ptr, ids = v.cells()
    ptr_ref, ids_ref = orc.voronoi_cells(cell_ref, c_ref, m)
    assert np.array_equal(ptr, ptr_ref) and np.array_equal(ids, ids_ref)
    v.close()
    ds.close()


@pytest.mark.parametrize("scale", [1e-22, 3e-3, 1e17])
def test_fps_screen_extreme_scales(orc, scale):
    """The BF16 screen of the farthest-point sweep is a proven bound: with
    coordinates so small that BF16 differences underflow when squared in
    FP32, or so large that FP32 squares approach overflow, and with integer
    data full of exact ties, the landmarks, cells and distances equal the
    exact oracle's."""
    l2, vor = _mods()
    rng = np.random.default_rng(5)
    pts = (rng.standard_normal((6000, 24)) * scale).astype(F)
    ints = rng.integers(0, 4, size=(3000, 8)).astype(F)
    for X, m in ((pts, 200), (ints, 40)):
        flush_print(f"fps screen scale={scale} n={X.shape[0]} d={X.shape[1]}")
        ds = l2.Dataset(X)
        v = vor.Voronoi(ds, fps=m, first=1)
        c_ref, cell_ref, dist_ref = orc.fps(X, m, 1)
        assert np.array_equal(v.centers(), c_ref)
        cell, dist = v.assignment()
        assert np.array_equal(cell, cell_ref)
        assert dist.tobytes() == dist_ref.tobytes()
        v.close()
        ds.close()


def _graph_case(orc, pts, centers, eps, check_ghosts=True):
    l2, vor = _mods()
    ds = l2.Dataset(pts)
    v = vor.Voronoi(ds, centers=centers)
    cell, dist = v.assignment()
    cell_ref, dist_ref = orc.voronoi_assign(pts, pts[centers])
    assert np.array_equal(cell, cell_ref)
    assert dist.tobytes() == dist_ref.tobytes()
    ptr, ids = v.cells()
    ptr_ref, ids_ref = orc.voronoi_cells(cell_ref, np.asarray(centers), len(centers))
    assert np.array_equal(ptr, ptr_ref) and np.array_equal(ids, ids_ref)
    g = v.graph(eps)
    ij, w = g.edges(sorted=True)
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    _same_edges(ij, w, ij_ref, w_ref)
    st = g.stats
    assert st.edges == ij_ref.shape[0] and st.band_flips <= st.band_pairs
    if check_ghosts:
        gh = v.ghosts()
        gh_ref = orc.ghosts(pts, cell_ref, dist_ref, pts[centers], eps)
        assert np.array_equal(gh, gh_ref)
        assert st.ghosts == gh_ref.shape[0] and st.ghost_candidates >= st.ghosts
    g.close()
    v.close()
    ds.close()
    return st


@pytest.mark.parametrize("scale", [1.0, 0.37])
def test_cells_at_the_adjacency_bound(orc, scale):
    """Two cells on a line, members at 0..R from their centers, whose closest
    cross pair is exactly eps apart: the centers are then R_a + R_b + eps
    apart, the largest separation the cell adjacency keeps, and that pair
    must be an edge (integer coordinates take the int8 filter, scaled ones
    the BF16 filter); every edge and ghost as the oracle's."""
    d = 16
    pts = np.zeros((82, d), np.float32)
    pts[:41, 0] = np.arange(-20, 21)
    pts[41:, 0] = np.arange(29, 70)
    pts[:, 1] = np.tile([0.0, 1.0], 41)    # a second coordinate: the pairs are not all collinear
    pts *= np.float32(scale)
    eps = float(orc.l2(pts[40], pts[41]))   # the closest cross pair, (20, 0) and (29, 1)
    centers = np.array([20, 61])            # x = 0 and x = 49
    flush_print(f"adjacency bound scale={scale} eps={eps}")
    st = _graph_case(orc, pts, centers, eps)
    assert st.edges > 0 and st.ghosts > 0


def test_gmm_small_landmark_graph(orc):
    from rnng_amd import synth
    _, vor = _mods()
    pts, eps = synth.points_gmm_small()
    centers = vor.selection_sample(pts.shape[0], 64, 1)
    flush_print(f"landmark gmm_small 10000x16 m=64 eps={eps}")
    st = _graph_case(orc, pts, centers, eps)
    assert st.edges > 50_000 and st.ghosts > 0


@pytest.mark.parametrize("n,d,m", [(6000, 96, 40), (4000, 128, 16), (2500, 33, 250)])
def test_random_landmark_graph(orc, n, d, m):
    rng = np.random.default_rng(n + d + m)
    k = 12
    mu = rng.standard_normal((k, d)).astype(F) * 3
    pts = (mu[rng.integers(0, k, n)] + rng.standard_normal((n, d)).astype(F)).astype(F)
    from rnng_amd import synth
    eps = synth.radius_for_degree(pts, 20.0)
    centers = np.sort(rng.choice(n, size=m, replace=False))
    flush_print(f"landmark random n={n} d={d} m={m} eps={eps}")
    _graph_case(orc, pts, centers, eps)


def test_integer_landmark_graph(orc):
    """Integer-valued data (the exact integer filter) with duplicate points."""
    from rnng_amd import synth
    rng = np.random.default_rng(11)
    pts, _ = synth.points_gmm_int(n=8000, d=128, k=16, seed=3)
    pts[4000:4100] = pts[:100]            # exact duplicates across cells
    eps = synth.radius_for_degree(pts, 25.0)
    centers = np.sort(rng.choice(pts.shape[0], size=50, replace=False))
    flush_print(f"landmark int 8000x128 m=50 eps={eps}")
    st = _graph_case(orc, pts, centers, eps)
    assert st.int8_filter == 1   # integer data spanning <= 255: the exact int8 cell filter
    from rnng_amd import _lib
    _lib.check(_lib.lib().rnng_set_int8_filter(0))
    try:
        st = _graph_case(orc, pts, centers, eps, check_ghosts=False)
        assert st.int8_filter == 0
    finally:
        _lib.check(_lib.lib().rnng_set_int8_filter(1))


@pytest.mark.parametrize("case", ["one_cell", "singletons", "eps_zero", "dup_centers", "huge_eps"])
def test_landmark_edge_cases(orc, case):
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 4, size=(400, 6)).astype(F)
    eps = F(1.5)
    if case == "huge_eps":
        # every pair is an edge and the threshold sits at its clamp: padding
        # positions between the cells must still never be candidates
        centers = np.arange(0, 400, 37)
        eps = F(1e6)
    elif case == "one_cell":
        centers = np.array([7])
    elif case == "singletons":
        centers = np.arange(400)
        pts = rng.standard_normal((400, 6)).astype(F)
        eps = F(1.2)
    elif case == "eps_zero":
        centers = np.arange(0, 400, 9)
        eps = F(0.0)
    else:
        # two centers at equal coordinates: the earlier one's cell is empty
        pts[10] = pts[20]
        centers = np.array([10, 20, 30, 40])
    flush_print(f"landmark edge case {case}")
    _graph_case(orc, pts, centers, eps)


def test_vmf_fps_landmark_graph(orc):
    """Unit-norm vMF-like data with farthest-point landmarks (the cell
    workload's shape at reduced size) against the all-pairs oracle."""
    from rnng_amd import synth
    l2, vor = _mods()
    pts, eps = synth.points_vmf(n=40000, d=96, k=64, seed=4)
    flush_print(f"vmf 40000x96 fps m=256 eps={eps}")
    ds = l2.Dataset(pts)
    v = vor.Voronoi(ds, fps=256, first=0)
    g = v.graph(eps)
    ij, w = g.edges(sorted=True)
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    _same_edges(ij, w, ij_ref, w_ref)
    flush_print(f"  edges={ij.shape[0]} ghosts={g.stats.ghosts} cand={g.stats.candidates} "
                f"pairs={g.stats.pair_evals}")
    g.close()
    v.close()
    ds.close()
