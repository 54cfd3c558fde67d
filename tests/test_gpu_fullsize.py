"""Full-size parity of the device paths against the oracle.

The oracle's brute force at these sizes takes from minutes to hours, so its
answers are frozen as digests (tests/golden/make_full_digests.py): edge count
plus SHA-256 of the sorted (i, j) int64 pairs and of the float32 weights (or
distances, or the (value, id)-sorted filtration).  The device output must give
the same digests.  The 2,000,000 x 96 cell workload is checked against the plain
kernel's all-pairs edge list at full size (ids and weights bit for bit) and its
ghosts against the reference's ghost criterion (include/point.hpp:894-918) in
the oracle's FP32 arithmetic."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fx(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return json.load(open(path))


def _sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


def test_gmm_int_1Mx128_all_paths_equal_oracle_digest():
    """The headline workload: the int8 filter (default on this integer data),
    the BF16 filter and the plain kernel all give the oracle's edge list."""
    from rnng_amd import Stats, _lib, l2, synth
    fx = _fx("full_gmm_int_1Mx128.json")
    pts, eps = synth.points_gmm_int()
    assert _sha(pts, np.float32) == fx["sha256_points_float32"]
    assert int(np.float32(eps).view(np.uint32)) == fx["eps_bits"]
    ds = l2.Dataset(pts)
    L = _lib.lib()
    try:
        for name, mode, int8 in (("int8 filter", "mfma", 1), ("bf16 filter", "mfma", 0), ("plain", "plain", 1)):
            flush_print(f"1M x 128 full size: {name}")
            _lib.check(L.rnng_set_int8_filter(int8))
            st = Stats()
            g = ds.graph(eps, mode, st)
            ij, w = g.edges(sorted=True)
            g.close()
            if name == "int8 filter":
                assert st.int8_filter == 1
            assert ij.shape[0] == fx["edges"], name
            assert _sha(ij, np.int64) == fx["sha256_ij_int64"], name
            assert _sha(w, np.float32) == fx["sha256_w_float32"], name
    finally:
        _lib.check(L.rnng_set_int8_filter(1))
        ds.close()


def test_edit_40k_subset_equals_oracle_digest():
    from rnng_amd import edit, synth
    fx = _fx("full_edit_40k.json")
    atoms, offs = synth.dna_mutants()
    n = fx["strings"]
    sub_atoms, sub_offs = atoms[: int(offs[n])], offs[: n + 1]
    assert _sha(sub_atoms, np.uint8) == fx["sha256_atoms"]
    flush_print(f"edit graph: first {n} of 200,000 strings, eps 8")
    S = edit.StringSet(sub_atoms, sub_offs)
    g = S.graph(8.0)
    ij, d = g.edges(sorted=True)
    g.close()
    S.close()
    assert ij.shape[0] == fx["edges"]
    assert _sha(ij, np.int64) == fx["sha256_ij_int64"]
    assert _sha(d, np.float32) == fx["sha256_dist_float32"]


def test_edit_200k_full_size_counts():
    """The full 200,000-string workload: the edge set contains the 40,000-string
    subset's edges (pairs with both ends in the subset) exactly."""
    from rnng_amd import edit, synth
    fx = _fx("full_edit_40k.json")
    atoms, offs = synth.dna_mutants()
    flush_print("edit graph: 200,000 strings, eps 8")
    S = edit.StringSet(atoms, offs)
    g = S.graph(8.0)
    ij, d = g.edges(sorted=True)
    g.close()
    S.close()
    n = fx["strings"]
    sub = (ij[:, 1] < n)
    assert int(sub.sum()) == fx["edges"]
    assert _sha(ij[sub], np.int64) == fx["sha256_ij_int64"]
    assert _sha(d[sub], np.float32) == fx["sha256_dist_float32"]
    assert (ij[:, 0] < ij[:, 1]).all() and ((d >= 0) & (d <= 8) & (d == np.rint(d))).all()


def test_rips_300kx3_filtration_equals_oracle_digest():
    from rnng_amd import l2, rips, synth
    fx = _fx("full_rips_300kx3.json")
    pts, eps = synth.points_rips()
    assert _sha(pts, np.float32) == fx["sha256_points_float32"]
    flush_print("rips 300k x 3, eps 0.08, dimension 2")
    ds = l2.Dataset(pts)
    with rips.Filtration(ds, eps, 2) as f:
        total, per = f.size()
        assert [int(x) for x in per[:3]] == fx["per_dim"]
        val, ids = f.copy()
    ds.close()
    assert _sha(val, np.float32) == fx["sha256_value_float32"]
    assert _sha(ids, np.int64) == fx["sha256_id_int64"]


def test_vmf_2Mx96_cells_equal_plain_and_ghosts_follow_the_reference(orc):
    """2,000,000 x 96 vMF points through 4,096 farthest-point cells with the
    shipped cost model and the autotuner: edges equal the plain kernel's
    all-pairs list bit for bit; ghosts satisfy the reference's criterion."""
    from rnng_amd import Stats, costmodel, l2, synth, voronoi
    fx_path = os.path.join(GOLD, "full_vmf_2Mx96.json")
    pts, eps = synth.points_vmf()
    assert costmodel.use_default()
    ds = l2.Dataset(pts)
    flush_print("vmf 2M x 96: 4,096 farthest-point cells")
    v = voronoi.Voronoi(ds, fps=4096, first=0)
    st = Stats()
    try:
        g = v.graph(eps, "mfma", st)
    finally:
        costmodel.set_cost_model(None)
    ij, w = g.edges(sorted=True)
    g.close()
    assert st.predicted_candidates > 0 and st.filter_passes == 1
    flush_print(f"cells: {ij.shape[0]} edges, {st.ghosts} ghosts; plain kernel next")
    gp = ds.graph(eps, "plain")
    ijp, wp = gp.edges(sorted=True)
    gp.close()
    assert ij.shape == ijp.shape and np.array_equal(ij, ijp)
    assert w.tobytes() == wp.tobytes()
    if os.path.exists(fx_path):
        fx = json.load(open(fx_path))
        assert ij.shape[0] == fx["edges"] and _sha(ij, np.int64) == fx["sha256_ij_int64"]
        assert _sha(w, np.float32) == fx["sha256_w_float32"]

    gh = v.ghosts()
    assert gh.shape[0] == st.ghosts > 0
    assert (np.diff(gh[:, 0] * (1 << 32) + gh[:, 1]) > 0).all()   # sorted, unique
    centers = v.centers()
    cell, delta = v.assignment()
    ptr, ids = v.cells()
    assert (cell[gh[:, 1]] != gh[:, 0]).all()
    rng = np.random.default_rng(0)
    f32 = np.float32
    for c, q in gh[rng.choice(gh.shape[0], size=400, replace=False)]:
        # L2(center_c, q) <= fl(delta_q + 2 eps) and a member of c within eps
        assert orc.l2(pts[centers[c]], pts[q]) <= f32(delta[q]) + f32(2) * f32(eps)
        members = ids[ptr[c]:ptr[c + 1]]
        assert orc.l2_eps_pairs_count(pts[q][None, :], eps, Y=pts[members]) > 0
    v.close()
    ds.close()


def test_vmf_200k_cells_with_the_reference_centers_equal_its_landmark_output():
    """The reference's landmark driver on the first 200,000 of the 2M x 96 vMF
    points (410 centers picked by its selection_sample with seed 7, 8 MPI
    ranks): the cell path with the same centers gives its sorted pairs bit for
    bit (and the oracle's brute force, when that digest is present)."""
    import json as _json
    from rnng_amd import l2, synth, voronoi
    ref = _json.load(open(os.path.join(GOLD, "ref_outputs.json"))).get("landmark_vmf_200kx96")
    if ref is None:
        pytest.skip("reference landmark run not recorded")
    pts, eps = synth.points_vmf()
    pts = np.ascontiguousarray(pts[:200_000])
    centers = voronoi.selection_sample(pts.shape[0], 410, 7)
    ds = l2.Dataset(pts)
    v = voronoi.Voronoi(ds, centers=centers)
    try:
        flush_print("vmf 200k x 96, 410 reference centers")
        g = v.graph(eps)
        ij, w = g.edges(sorted=True)
        g.close()
        assert ij.shape[0] == ref["edges_undirected"]
        assert _sha(ij, np.int64) == ref["sha256_pairs_int64"]
        path = os.path.join(GOLD, "full_vmf_200kx96.json")
        if os.path.exists(path):
            fx = _json.load(open(path))
            assert _sha(ij, np.int64) == fx["sha256_ij_int64"] and _sha(w, np.float32) == fx["sha256_w_float32"]
    finally:
        v.close()
        ds.close()
