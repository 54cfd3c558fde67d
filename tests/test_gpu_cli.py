"""The command-line drivers end to end on the GPU: input files in the
reference's formats, output files in the reference's formats, contents equal
to what the oracle's pairs and filtration give when formatted as the
reference's writers format them (include/graph.hpp:20-57 through one rank's
redistribute_edges, include/simplex.hpp:175-192)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "rnng_amd", "bin")


def _write_fvecs(path, X):
    with open(path, "wb") as f:
        for row in np.asarray(X, np.float32):
            f.write(struct.pack("<i", row.shape[0]))
            f.write(row.tobytes())


def _expected_graph_lines(n, ij, w):
    """One rank of the drivers: both orientations of every edge, a zero loop
    per vertex (each point finds itself), sorted, written with setprecision(6)."""
    u = np.concatenate([ij[:, 0], ij[:, 1], np.arange(n)])
    v = np.concatenate([ij[:, 1], ij[:, 0], np.arange(n)])
    ww = np.concatenate([w, w, np.zeros(n, np.float32)])
    o = np.lexsort((v, u))
    lines = [f"% {n} {n} {o.size}"]
    lines += ["%d %d %s" % (a, b, "%.6g" % float(c)) for a, b, c in zip(u[o], v[o], ww[o])]
    return lines


def _run(args):
    flush_print(" ".join(os.path.basename(a) for a in args))
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_systolic_l2_file(orc, tmp_path):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_small(n=6000)
    f = tmp_path / "p.fvecs"
    _write_fvecs(f, pts)
    out = tmp_path / "g.txt"
    _run([os.path.join(BIN, "rnng_systolic"), "-i", str(f), "-r", repr(float(eps)), "-o", str(out)])
    ij, w = orc.l2_eps_pairs(pts, eps)
    assert out.read_text().splitlines() == _expected_graph_lines(pts.shape[0], ij, w)


def test_systolic_l2_file_3d(orc, tmp_path):
    """3-D points: the driver's self join runs on the uniform grid."""
    pts = np.random.default_rng(8).uniform(0, 4, size=(5000, 3)).astype(np.float32)
    eps = np.float32(0.2)
    f = tmp_path / "p.fvecs"
    _write_fvecs(f, pts)
    out = tmp_path / "g.txt"
    _run([os.path.join(BIN, "rnng_systolic"), "-i", str(f), "-r", repr(float(eps)), "-o", str(out)])
    ij, w = orc.l2_eps_pairs(pts, eps)
    assert ij.shape[0] > 1000
    assert out.read_text().splitlines() == _expected_graph_lines(pts.shape[0], ij, w)


def test_systolic_fbin_and_u8bin_inputs(orc, tmp_path):
    """Inputs named .fbin / .u8bin are read by their header (dataset_io.py
    formats); the graph is the oracle's on the same points."""
    from rnng_amd import dataset_io
    X = np.random.default_rng(9).integers(0, 40, size=(3000, 6)).astype(np.uint8)
    eps = np.float32(14.5)
    ij, w = orc.l2_eps_pairs(X.astype(np.float32), eps)
    for ext in (".fbin", ".u8bin"):
        f = str(tmp_path / ("p" + ext))
        dataset_io.write_file(f, X if ext == ".u8bin" else X.astype(np.float32))
        out = tmp_path / ("g" + ext + ".txt")
        _run([os.path.join(BIN, "rnng_systolic"), "-i", f, "-r", repr(float(eps)), "-o", str(out)])
        assert out.read_text().splitlines() == _expected_graph_lines(X.shape[0], ij, w)


def test_systolic_edit_file(orc, tmp_path):
    from rnng_amd import synth
    atoms, offs = synth.dna_mutants(n_seeds=40, per_seed=25)
    f = tmp_path / "s.seqs"
    with open(f, "wb") as fh:
        for i in range(offs.shape[0] - 1):
            fh.write(atoms[offs[i]:offs[i + 1]].tobytes() + b"\n")
    out = tmp_path / "g.txt"
    _run([os.path.join(BIN, "rnng_systolic"), "-D", "edit", "-i", str(f), "-r", "8", "-o", str(out)])
    ij, d = orc.edit_eps_pairs(atoms, offs, 8)
    assert out.read_text().splitlines() == _expected_graph_lines(offs.shape[0] - 1, ij, d.astype(np.float32))


def test_landmark_file_equals_systolic(orc, tmp_path):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_small(n=8000)
    f = tmp_path / "p.fvecs"
    _write_fvecs(f, pts)
    out = tmp_path / "g.txt"
    r = _run([os.path.join(BIN, "rnng_landmark"), "-i", str(f), "-r", repr(float(eps)), "-m", "48", "-s", "3",
              "-o", str(out)])
    assert "added ghost points" in r.stderr
    ij, w = orc.l2_eps_pairs(pts, eps)
    assert out.read_text().splitlines() == _expected_graph_lines(pts.shape[0], ij, w)


@pytest.mark.parametrize("use_ids", [True, False])
def test_rips_file(orc, tmp_path, use_ids):
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 1, size=(1500, 3)).astype(np.float32)
    eps = np.float32(0.09)
    f = tmp_path / "p.fvecs"
    _write_fvecs(f, pts)
    prefix = tmp_path / "filt"
    args = [os.path.join(BIN, "rnng_rips"), "-i", str(f), "-r", repr(float(eps)), "-d", "2", "-o", str(prefix)]
    if not use_ids:
        args.append("-V")
    _run(args)
    val, ids = orc.rips_filtration(pts, eps, 2)
    if use_ids:
        want = ["%f\t%d" % (float(v), int(i)) for v, i in zip(val, ids)]
    else:
        from itertools import combinations
        # vertex lists from the ids by brute force over the edges' endpoints
        ij, _ = orc.l2_eps_pairs(pts, eps)
        rep = {}
        for v in range(pts.shape[0]):
            rep[orc.simplex_id(np.array([v]))] = f"<{v}>"
        for a, b in ij:
            rep[orc.simplex_id(np.array([a, b]))] = f"<{a},{b}>"
        adj = {}
        for a, b in ij:
            adj.setdefault(int(a), set()).add(int(b))
            adj.setdefault(int(b), set()).add(int(a))
        for a, nb in adj.items():
            for b, c in combinations(sorted(x for x in nb if x > a), 2):
                if c in adj.get(b, ()):
                    rep[orc.simplex_id(np.array([a, b, c]))] = f"<{a},{b},{c}>"
        want = ["%f\t%s" % (float(v), rep[int(i)]) for v, i in zip(val, ids)]
    assert (tmp_path / "filt.txt").read_text().splitlines() == want


def test_rips_default_dimension_three(orc, tmp_path):
    """Without -d the driver builds up to tetrahedra (the reference's default)."""
    rng = np.random.default_rng(8)
    pts = rng.uniform(0, 1, size=(600, 3)).astype(np.float32)
    eps = np.float32(0.16)
    f = tmp_path / "p.fvecs"
    _write_fvecs(f, pts)
    prefix = tmp_path / "filt"
    _run([os.path.join(BIN, "rnng_rips"), "-i", str(f), "-r", repr(float(eps)), "-o", str(prefix)])
    val, ids = orc.rips_filtration(pts, eps, 3)
    assert (tmp_path / "filt.txt").read_text().splitlines() == ["%f\t%d" % (float(v), int(i)) for v, i in zip(val, ids)]
    r = subprocess.run([os.path.join(BIN, "rnng_rips"), "-i", str(f), "-r", "0.5", "-d", "4"], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "dimension 3" in r.stderr
