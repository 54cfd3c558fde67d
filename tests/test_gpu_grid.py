"""GPU parity of the uniform-grid self join (RNNG_MODE_GRID, d <= 3) against
the CPU oracle's brute-force pairs: the same sorted (i, j) list and the same
float weights bit for bit, on spread, clustered, surface and lattice inputs
(lattice distances land exactly on eps), and on the edge cases of the
reference's semantics (duplicates at eps = 0, non-finite coordinates, a grid
that has to grow its cells)."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu
F = np.float32


def _l2():
    from rnng_amd import l2
    return l2


def _check_same(ij, w, ij_ref, w_ref):
    assert ij.shape == ij_ref.shape, (ij.shape, ij_ref.shape)
    assert np.array_equal(ij, ij_ref)
    assert w.tobytes() == w_ref.tobytes()


def _lattice(k, d):
    ax = [np.arange(k, dtype=F)] * d
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, d).astype(F)


@pytest.mark.parametrize("kind,n,d,eps", [
    ("normal", 5000, 1, 0.002),
    ("normal", 6000, 2, 0.05),
    ("normal", 8000, 3, 0.25),
    ("clusters", 7000, 3, 0.08),
    ("clusters", 4000, 2, 0.02),
    ("lattice", 0, 3, 1.0),
    ("lattice", 0, 2, 1.5),
    ("lattice", 0, 1, 2.0),
])
def test_grid_matches_oracle(orc, kind, n, d, eps):
    from rnng_amd import Stats
    rng = np.random.default_rng(n + 10 * d)
    if kind == "normal":
        X = rng.standard_normal((n, d)).astype(F)
    elif kind == "clusters":
        c = rng.uniform(-5, 5, size=(20, d))
        X = (c[rng.integers(0, 20, n)] + 0.05 * rng.standard_normal((n, d))).astype(F)
    else:
        X = _lattice({1: 3000, 2: 60, 3: 16}[d], d)
        X = X[rng.permutation(X.shape[0])]
        n = X.shape[0]
    eps = F(eps)
    flush_print(f"grid {kind} n={n} d={d} eps={eps}")
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    st = Stats()
    ij, w = _l2().eps_pairs(X, eps, "grid", stats=st)
    _check_same(ij, w, ij_ref, w_ref)
    assert ij.shape[0] > 0
    assert st.edges == ij.shape[0]
    assert st.edges <= st.pair_evals < n * (n - 1) // 2


def test_grid_rips_surfaces(orc):
    from rnng_amd import synth
    pts, eps = synth.points_rips(n=20000)
    flush_print(f"grid rips surfaces 20000x3 eps={eps}")
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    ij, w = _l2().eps_pairs(pts, eps, "grid")
    _check_same(ij, w, ij_ref, w_ref)


def test_grid_edge_cases(orc):
    l2 = _l2()
    rng = np.random.default_rng(9)
    X = rng.standard_normal((400, 3)).astype(F)
    X[7] = X[3]
    X[250] = X[3]
    flush_print("grid edge cases")
    ij, w = l2.eps_pairs(X, 0.0, "grid")   # eps = 0: exact duplicates only
    assert [tuple(p) for p in ij] == [(3, 7), (3, 250), (7, 250)]
    assert np.all(w == 0)
    assert l2.eps_pairs(X, -1.0, "grid")[0].shape[0] == 0
    assert l2.eps_pairs(X, float("nan"), "grid")[0].shape[0] == 0
    for m in (0, 1):
        assert l2.eps_pairs(X[:m], 1.0, "grid")[0].shape[0] == 0
    ij, w = l2.eps_pairs(X[:2], 100.0, "grid")
    assert [tuple(p) for p in ij] == [(0, 1)]
    # infinite eps: every pair (the plain kernel runs; nothing to prune)
    ij, w = l2.eps_pairs(X[:100], float("inf"), "grid")
    assert ij.shape[0] == 100 * 99 // 2
    # all points equal: one cell
    Z = np.full((300, 2), 3.5, F)
    assert l2.eps_pairs(Z, 0.0, "grid")[0].shape[0] == 300 * 299 // 2


def test_grid_nonfinite_coordinates(orc):
    rng = np.random.default_rng(11)
    X = rng.standard_normal((3000, 3)).astype(F)
    X[5, 0] = np.nan
    X[17, 2] = np.inf
    X[18, 2] = np.inf
    X[40, 1] = -np.inf
    eps = F(0.3)
    flush_print("grid non-finite coordinates")
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    ij, w = _l2().eps_pairs(X, eps, "grid")
    _check_same(ij, w, ij_ref, w_ref)
    assert not np.isin([5, 17, 18, 40], ij).any()


def test_grid_growth_and_tiny_eps(orc):
    """A bounding box of 1e30 with eps 1e-3 cannot be a dense grid of side
    eps: the cells grow, the edges stay the same.  And a tiny eps with
    near-duplicates 1e-31 apart."""
    rng = np.random.default_rng(12)
    far = rng.uniform(-1e30, 1e30, size=(200, 3))
    near = 1e-3 * rng.standard_normal((2000, 3))
    X = np.concatenate([far, near]).astype(F)
    X = X[rng.permutation(X.shape[0])]
    eps = F(1.5e-3)
    flush_print("grid growth")
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    ij, w = _l2().eps_pairs(X, eps, "grid")
    _check_same(ij, w, ij_ref, w_ref)
    assert ij.shape[0] > 1000
    Y = rng.standard_normal((500, 2)).astype(F)
    Y = np.concatenate([Y, Y + F(1e-31)]).astype(F)
    eps = F(1e-30)
    ij_ref, w_ref = orc.l2_eps_pairs(Y, eps)
    ij, w = _l2().eps_pairs(Y, eps, "grid")
    _check_same(ij, w, ij_ref, w_ref)


def test_grid_unsupported_shapes():
    from rnng_amd._lib import RNNG_ERR_UNSUPPORTED, RnngError
    l2 = _l2()
    X = np.zeros((10, 4), F)
    with pytest.raises(RnngError) as ei:
        l2.eps_pairs(X, 1.0, "grid")
    assert ei.value.code == RNNG_ERR_UNSUPPORTED
    with pytest.raises(RnngError) as ei:
        l2.eps_bipartite(np.zeros((5, 3), F), np.zeros((6, 3), F), 1.0, "grid")
    assert ei.value.code == RNNG_ERR_UNSUPPORTED


def test_grid_oneshot_entry(orc):
    rng = np.random.default_rng(13)
    X = rng.uniform(0, 10, size=(5000, 3)).astype(F)
    eps = F(0.4)
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    ij, w = _l2().eps_pairs_oneshot(X, eps, "grid")
    _check_same(ij, w, ij_ref, w_ref)


def test_grid_full_rips_size_equals_plain():
    """300,000 x 3 (torus, sphere, box) at eps 0.08: the grid's edge list is
    the plain all-pairs kernel's, weights bit for bit."""
    from rnng_amd import Stats, l2, synth
    pts, eps = synth.points_rips()
    flush_print(f"grid vs plain 300000x3 eps={eps}")
    ds = l2.Dataset(pts)
    st = Stats()
    a = ds.graph(eps, "grid", st)
    b = ds.graph(eps, "plain")
    ij_a, w_a = a.edges(sorted=True)
    ij_b, w_b = b.edges(sorted=True)
    _check_same(ij_a, w_a, ij_b, w_b)
    flush_print(f"edges {ij_a.shape[0]} grid stats {st.as_dict()}")
    assert st.pair_evals < 300_000 * 299_999 // 2 // 100
    a.close()
    b.close()
    ds.close()
