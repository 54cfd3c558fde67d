"""Checks against the committed golden fixtures (tests/golden/, written by
tests/golden/make_fixtures.py from closed forms, textbook values and the
10,000 x 16 workload).  The CPU tests pin the oracle, the host functions and
the generators to the fixtures; the GPU tests pin the device paths to them."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import flush_print

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


def test_l2_known_pairs_oracle(orc):
    for p in _load("l2_known.json")["pairs"]:
        a = np.array(p["a"], np.float32)
        b = np.array(p["b"], np.float32)
        assert orc.l2sq(a, b) == np.float32(p["sq"])
        assert orc.l2(a, b) == np.sqrt(np.float32(p["sq"]))


def test_thresholds_oracle_and_library(orc):
    from rnng_amd import l2
    for t in _load("l2_known.json")["thresholds"]:
        eps = np.array([t["eps_bits"]], np.uint32).view(np.float32)[0]
        want = np.array([t["t_bits"]], np.uint32).view(np.float32)[0]
        assert orc.sq_threshold(eps) == want
        assert l2.sq_threshold(eps) == want


def test_edit_known_oracle(orc):
    for e in _load("edit_known.json"):
        assert orc.edit_distance(e["s"].encode(), e["t"].encode()) == e["d"]


def test_selection_fixtures(orc):
    from rnng_amd import voronoi
    for s in _load("selection.json"):
        assert orc.selection_sample(s["range"], s["size"], s["seed"]).tolist() == s["sample"]
        assert voronoi.selection_sample(s["range"], s["size"], s["seed"]).tolist() == s["sample"]


def test_simplex_id_fixtures(orc):
    for s in _load("simplex_ids.json"):
        assert orc.simplex_id(np.array(s["verts"])) == s["id"]


def test_gmm_small_generator_and_oracle(orc):
    from rnng_amd import synth
    g = _load("gmm_small.json")
    pts, eps = synth.points_gmm_small()
    assert _sha(pts, np.float32) == g["sha256_points_float32"]
    assert int(np.float32(eps).view(np.uint32)) == g["eps_bits"]
    ij, w = orc.l2_eps_pairs(pts, eps)
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(w, np.float32) == g["sha256_w_float32"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["plain", "mfma", "cells"])
def test_gmm_small_device_paths_match_fixture(path):
    from rnng_amd import l2, synth, voronoi
    g = _load("gmm_small.json")
    pts, _ = synth.points_gmm_small()
    eps = np.array([g["eps_bits"]], np.uint32).view(np.float32)[0]
    flush_print(f"golden gmm_small via {path}")
    if path == "cells":
        ds = l2.Dataset(pts)
        v = voronoi.Voronoi(ds, centers=voronoi.selection_sample(pts.shape[0], 100, 12))
        gr = v.graph(eps)
        ij, w = gr.edges(sorted=True)
        gr.close()
        v.close()
        ds.close()
    else:
        ij, w = l2.eps_pairs(pts, eps, path)
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(w, np.float32) == g["sha256_w_float32"]


def test_gmm_int_small_generator_and_oracle(orc):
    from rnng_amd import synth
    g = _load("gmm_int_small.json")
    pts, eps = synth.points_gmm_int(n=g["n"])
    assert _sha(pts, np.float32) == g["sha256_points_float32"]
    assert int(np.float32(eps).view(np.uint32)) == g["eps_bits"]
    ij, w = orc.l2_eps_pairs(pts, eps)
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(w, np.float32) == g["sha256_w_float32"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["plain", "mfma", "mfma_bf16", "cells"])
def test_gmm_int_small_device_paths_match_fixture(path):
    """Integer-valued data: the MFMA and cell paths take the exact int8 filter
    (and the exact integer recheck); with the int8 filter switched off they
    take BF16; every path gives the frozen edge set."""
    from rnng_amd import Stats, _lib, l2, synth, voronoi
    g = _load("gmm_int_small.json")
    pts, _ = synth.points_gmm_int(n=g["n"])
    eps = np.array([g["eps_bits"]], np.uint32).view(np.float32)[0]
    flush_print(f"golden gmm_int_small via {path}")
    st = Stats()
    if path == "cells":
        ds = l2.Dataset(pts)
        v = voronoi.Voronoi(ds, fps=64, first=0)
        gr = v.graph(eps)
        ij, w = gr.edges(sorted=True)
        st = gr.stats
        gr.close()
        v.close()
        ds.close()
    elif path == "mfma_bf16":
        _lib.check(_lib.lib().rnng_set_int8_filter(0))
        try:
            ij, w = l2.eps_pairs(pts, eps, "mfma", stats=st)
        finally:
            _lib.check(_lib.lib().rnng_set_int8_filter(1))
    else:
        ij, w = l2.eps_pairs(pts, eps, path, stats=st)
    assert st.int8_filter == (1 if path in ("mfma", "cells") else 0)
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(w, np.float32) == g["sha256_w_float32"]


@pytest.mark.gpu
def test_l2_known_pairs_device():
    from rnng_amd import l2
    for p in _load("l2_known.json")["pairs"]:
        a = np.array([p["a"]], np.float32)
        b = np.array([p["b"]], np.float32)
        eps = np.float32(np.sqrt(np.float32(p["sq"])))
        for mode in ("plain", "mfma"):
            ij, w = l2.eps_bipartite(a, b, eps, mode)
            assert ij.tolist() == [[0, 0]] and w[0] == eps
            ij, _ = l2.eps_bipartite(a, b, np.nextafter(eps, np.float32(0)), mode)
            assert ij.shape[0] == (1 if p["sq"] == 0 else 0)


def test_edit_small_oracle(orc):
    from rnng_amd import synth
    g = _load("edit_small.json")
    atoms, offs = synth.dna_mutants(n_seeds=g["n_seeds"], per_seed=g["per_seed"])
    assert hashlib.sha256(np.ascontiguousarray(atoms, np.uint8).tobytes()).hexdigest() == g["sha256_atoms"]
    ij, d = orc.edit_eps_pairs(atoms, offs, g["eps"])
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(d, np.float32) == g["sha256_dist_float32"]


@pytest.mark.gpu
def test_edit_small_device_matches_fixture():
    from rnng_amd import edit, synth
    g = _load("edit_small.json")
    atoms, offs = synth.dna_mutants(n_seeds=g["n_seeds"], per_seed=g["per_seed"])
    ij, d = edit.eps_pairs(atoms, offs, g["eps"])
    assert ij.shape[0] == g["edges"]
    assert _sha(ij, np.int64) == g["sha256_ij_int64"] and _sha(d, np.float32) == g["sha256_dist_float32"]


def _rips_case(r):
    from rnng_amd import synth
    pts, _ = synth.points_rips(n=r["n"])
    assert _sha(pts, np.float32) == r["sha256_points_float32"]
    return pts, np.uint32(r["eps_bits"]).view(np.float32)


def test_rips_small_oracle(orc):
    for r in _load("rips_small.json"):
        pts, eps = _rips_case(r)
        val, ids = orc.rips_filtration(pts, eps, r["maxdim"])
        dims = (ids >> 60) & 0xF
        assert [int((dims == p).sum()) for p in range(r["maxdim"] + 1)] == r["per_dim"]
        assert _sha(val, np.float32) == r["sha256_value_float32"] and _sha(ids, np.int64) == r["sha256_id_int64"]


@pytest.mark.gpu
def test_rips_small_device_matches_fixture():
    from rnng_amd import rips
    for r in _load("rips_small.json"):
        pts, eps = _rips_case(r)
        val, ids = rips.rips_filtration(pts, eps, r["maxdim"])
        assert _sha(val, np.float32) == r["sha256_value_float32"] and _sha(ids, np.int64) == r["sha256_id_int64"]
