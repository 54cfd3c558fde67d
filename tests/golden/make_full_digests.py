"""Writes the oracle's digests of the full-size workloads (run from the
repository root: python tests/golden/make_full_digests.py [name ...]).

The oracle's brute force takes tens of minutes to hours at these sizes, so it
runs once here and the digests are frozen as fixtures; the GPU tests compare
the device output against them (edge count plus SHA-256 of the sorted pairs and
of the weights), which pins full-size completeness against the oracle rather
than against another device path.

  full_gmm_int_1Mx128.json   1,000,000 x 128 integer-valued GMM, the headline
                             self join (sorted (i, j) int64 pairs, float32 weights)
  full_rips_300kx3.json      300,000 x 3 torus/sphere/box points, eps 0.08,
                             maxdim 2: the (value, id)-sorted filtration
  full_edit_40k.json         the first 40,000 strings of the 200,000-string DNA
                             set, eps 8: sorted pairs and int32 distances
  full_vmf_2Mx96.json        2,000,000 x 96 vMF unit vectors: sorted pairs and weights
  full_vmf_200kx96.json      the first 200,000 of those points at the same eps (the
                             reference's landmark driver was run on this subset)

Each file also holds the SHA-256 of the generated input, so a change of the
generator shows up as a fixture mismatch instead of a parity failure.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def _sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


def _dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)
    print(f"{name} written: {obj}", flush=True)


def gmm_int():
    import oracle
    from rnng_amd import synth
    pts, eps = synth.points_gmm_int()
    t = time.time()
    ij, w = oracle.l2_eps_pairs(pts, eps, cap=13_000_000)
    _dump("full_gmm_int_1Mx128.json", {
        "n": int(pts.shape[0]), "d": int(pts.shape[1]), "eps_bits": int(np.float32(eps).view(np.uint32)),
        "edges": int(ij.shape[0]), "sha256_ij_int64": _sha(ij, np.int64), "sha256_w_float32": _sha(w, np.float32),
        "sha256_points_float32": _sha(pts, np.float32), "oracle_s": round(time.time() - t, 1)})


def vmf():
    import oracle
    from rnng_amd import synth
    pts, eps = synth.points_vmf()
    t = time.time()
    ij, w = oracle.l2_eps_pairs(pts, eps, cap=17_000_000)
    _dump("full_vmf_2Mx96.json", {
        "n": int(pts.shape[0]), "d": int(pts.shape[1]), "eps_bits": int(np.float32(eps).view(np.uint32)),
        "edges": int(ij.shape[0]), "sha256_ij_int64": _sha(ij, np.int64), "sha256_w_float32": _sha(w, np.float32),
        "sha256_points_float32": _sha(pts, np.float32), "oracle_s": round(time.time() - t, 1)})


def vmf200k():
    import oracle
    from rnng_amd import synth
    pts, eps = synth.points_vmf()
    pts = np.ascontiguousarray(pts[:200_000])
    t = time.time()
    ij, w = oracle.l2_eps_pairs(pts, eps, cap=2_000_000)
    _dump("full_vmf_200kx96.json", {
        "n": int(pts.shape[0]), "d": int(pts.shape[1]), "eps_bits": int(np.float32(eps).view(np.uint32)),
        "edges": int(ij.shape[0]), "sha256_ij_int64": _sha(ij, np.int64), "sha256_w_float32": _sha(w, np.float32),
        "sha256_points_float32": _sha(pts, np.float32), "oracle_s": round(time.time() - t, 1)})


def rips():
    import oracle
    from rnng_amd import synth
    pts, eps = synth.points_rips()
    t = time.time()
    val, ids = oracle.rips_filtration(pts, eps, 2, cap=175_000_000)
    dims = (ids >> 60) & 0xF
    _dump("full_rips_300kx3.json", {
        "n": int(pts.shape[0]), "maxdim": 2, "eps_bits": int(np.float32(eps).view(np.uint32)),
        "per_dim": [int((dims == p).sum()) for p in range(3)],
        "sha256_value_float32": _sha(val, np.float32), "sha256_id_int64": _sha(ids, np.int64),
        "sha256_points_float32": _sha(pts, np.float32), "oracle_s": round(time.time() - t, 1)})


def edit():
    import oracle
    from rnng_amd import synth
    atoms, offs = synth.dna_mutants()
    n = 40_000
    sub_atoms = atoms[: int(offs[n])]
    sub_offs = offs[: n + 1]
    t = time.time()
    ij, dist = oracle.edit_eps_pairs(sub_atoms, sub_offs, 8, cap=2_000_000)
    _dump("full_edit_40k.json", {
        "strings": n, "of": int(offs.shape[0] - 1), "eps": 8, "edges": int(ij.shape[0]),
        "sha256_ij_int64": _sha(ij, np.int64), "sha256_dist_float32": _sha(dist.astype(np.float32), np.float32),
        "sha256_atoms": _sha(sub_atoms, np.uint8), "oracle_s": round(time.time() - t, 1)})


JOBS = {"gmm_int": gmm_int, "rips": rips, "edit": edit, "vmf": vmf, "vmf200k": vmf200k}

if __name__ == "__main__":
    for name in (sys.argv[1:] or ["gmm_int", "rips", "edit"]):
        JOBS[name]()
