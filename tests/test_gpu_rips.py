"""GPU parity of the Rips filtration (vertices, edges, triangles,
tetrahedra) against the CPU oracle: identical (value, id) sequences, bit for
bit."""
import math

import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _rips():
    from rnng_amd import rips
    return rips


@pytest.mark.parametrize("maxdim", [0, 1, 2, 3])
def test_rips_small_bit_exact(orc, maxdim):
    from rnng_amd import synth
    pts, eps = synth.points_rips(n=3000)
    flush_print(f"rips 3000 eps={eps} maxdim={maxdim}")
    v_ref, i_ref = orc.rips_filtration(pts, eps, maxdim=maxdim)
    v, i = _rips().rips_filtration(pts, eps, maxdim)
    assert v.tobytes() == v_ref.tobytes()
    assert np.array_equal(i, i_ref)


@pytest.mark.parametrize("n,d,eps", [(500, 3, 0.3), (400, 5, 0.9), (300, 17, 3.0)])
def test_rips_random_bit_exact(orc, n, d, eps):
    rng = np.random.default_rng(n + d)
    pts = rng.standard_normal((n, d)).astype(np.float32)
    flush_print(f"rips random n={n} d={d}")
    v_ref, i_ref = orc.rips_filtration(pts, eps, maxdim=2)
    v, i = _rips().rips_filtration(pts, eps, 2)
    assert v.tobytes() == v_ref.tobytes()
    assert np.array_equal(i, i_ref)


@pytest.mark.parametrize("n,d,eps", [(400, 3, 0.45), (250, 6, 1.6), (2000, 3, 0.35)])
def test_rips_tetrahedra_bit_exact(orc, n, d, eps):
    rng = np.random.default_rng(n * d)
    pts = rng.standard_normal((n, d)).astype(np.float32)
    flush_print(f"rips dim 3 n={n} d={d}")
    v_ref, i_ref = orc.rips_filtration(pts, eps, maxdim=3)
    v, i = _rips().rips_filtration(pts, eps, 3)
    assert v.tobytes() == v_ref.tobytes()
    assert np.array_equal(i, i_ref)
    assert (_rips().simplex_dims(i) == 3).sum() > 0


@pytest.mark.parametrize("maxdim", [1, 2, 3])
def test_rips_value_ties_bit_exact(orc, maxdim):
    """A 5x5x5 lattice with some points duplicated: many simplices share a
    value (1, sqrt 2), edges of weight 0 share the vertices' value group, and
    lattice distances sit exactly at eps = 1 -- the order inside groups of
    equal value is (dimension, id)."""
    g = np.stack(np.meshgrid(*[np.arange(5, dtype=np.float32)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = np.concatenate([g, g[::7]]).astype(np.float32)
    pts = pts[np.random.default_rng(3).permutation(pts.shape[0])]
    for eps in (1.0, 1.5):
        flush_print(f"rips lattice ties eps={eps} maxdim={maxdim}")
        v_ref, i_ref = orc.rips_filtration(pts, eps, maxdim=maxdim)
        v, i = _rips().rips_filtration(pts, eps, maxdim)
        assert v.tobytes() == v_ref.tobytes()
        assert np.array_equal(i, i_ref)
        assert ((v == 0) & (_rips().simplex_dims(i) == 1)).sum() >= 18


def test_rips_clique_counts():
    pts = (np.random.default_rng(1).standard_normal((30, 3)) * 1e-3).astype(np.float32)
    v, i = _rips().rips_filtration(pts, 1.0, 3)
    dims = _rips().simplex_dims(i)
    assert [(dims == p).sum() for p in range(4)] == [30, math.comb(30, 2), math.comb(30, 3), math.comb(30, 4)]


def test_rips_dimension_three_size_guard():
    """Above 55,111 vertices the reference's int64 binomial for 4-subsets
    overflows: the library refuses instead of writing wrong ids."""
    from rnng_amd._lib import RNNG_ERR_UNSUPPORTED, RnngError
    pts = np.zeros((55_112, 3), np.float32)
    with pytest.raises(RnngError) as ei:
        _rips().rips_filtration(pts, 0.0, 3)
    assert ei.value.code == RNNG_ERR_UNSUPPORTED


def test_rips_full_size_properties():
    """300,000 x 3 at eps 0.08: sorted by (value, id), ids unique, counts per
    dimension consistent with the epsilon graph, every triangle's value is the
    max of its three edges' values."""
    from rnng_amd import Stats, l2, synth
    pts, eps = synth.points_rips()
    st = Stats()
    flush_print("rips 300000x3 eps=0.08")
    v, i = _rips().rips_filtration(pts, eps, 2, st)
    flush_print(f"simplices {v.shape[0]} stats {st.as_dict()}")
    dims = _rips().simplex_dims(i)
    assert np.all(np.diff(v) >= 0)
    same = np.diff(v) == 0
    assert np.all(np.diff(i)[same] > 0)
    assert np.unique(i).shape[0] == i.shape[0]
    assert (dims == 0).sum() == pts.shape[0]
    ne = (dims == 1).sum()
    g = l2.Dataset(pts).graph(eps, "plain")
    assert g.num_edges == ne
    # triangles: decode a sample and check the value
    import oracle
    tri = np.nonzero(dims == 2)[0]
    assert tri.shape[0] > 1_000_000
    rng = np.random.default_rng(0)
    for t in rng.choice(tri, 100, replace=False):
        uid = int(i[t]) & ((1 << 60) - 1)
        # invert the combinatorial number system
        w = int(round((6 * uid) ** (1 / 3))) + 2
        while math.comb(w, 3) > uid:
            w -= 1
        r = uid - math.comb(w, 3)
        vv = int((2 * r) ** 0.5) + 1
        while math.comb(vv, 2) > r:
            vv -= 1
        u = r - math.comb(vv, 2)
        assert u < vv < w
        e = [oracle.l2(pts[a], pts[b]) for a, b in ((u, vv), (u, w), (vv, w))]
        assert max(e) <= eps and np.float32(max(e)) == v[t]
