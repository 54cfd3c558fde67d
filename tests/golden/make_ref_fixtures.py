"""Runs the reference's own drivers (built from their sources by
`make -C oracle ref` into oracle/_ref) on this repository's seeded inputs and
freezes what they wrote as fixtures in tests/golden/ref_outputs.json (run from
the repository root: python tests/golden/make_ref_fixtures.py [job ...]).

Each entry names the driver, its flags and MPI ranks, the SHA-256 of the input
file it read and of the output file it wrote, the output's size and counts,
and the driver's own phase timers.  The tests compare this build's command-line
drivers and the oracle against these digests: the same input file gives the
same output file, byte for byte.

Jobs:
  c1          systolic and landmark drivers, 10,000 x 16 GMM (.fvecs)
  edit600     systolic and landmark drivers, -D edit, 600 DNA-like strings, eps 8
              (the landmark driver with 24 centers, two of them the same string,
              and with 3 distinct centers)
  rips        rips driver: 5,000 torus/sphere/box points (eps 0.3, -d 2, ids)
              and 1,500 points (eps 0.35, -d 3, vertex lists -V)
  edit40k     systolic driver, -D edit, the first 40,000 of the 200,000 strings
  gmm_int     systolic driver, the 1,000,000 x 128 integer-valued GMM (full size)
  c2scale     systolic driver on the first 50,000 / 100,000 / 200,000 points of that
              GMM at its eps (timing: how the cover trees scale)
  c4sub       landmark driver on the first 200,000 of the 2,000,000 x 96 vMF points at
              their eps, 410 random centers (-F; the full workload has 4,096)
  c5cells     rips driver on the 300,000 x 3 torus/sphere/box points, eps 0.08, -d 2,
              256 cells (-m 32 per rank; one cell of 300,000 points is out of reach)

The reference writes its output through MPI-IO; MPICH comes from MPI_HOME
(default /opt/conda).  Nothing here is imported by the product.
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.path.join(ROOT, "oracle", "_ref")
MPI_HOME = os.environ.get("MPI_HOME", "/opt/conda")
OUT = os.path.join(HERE, "ref_outputs.json")


def _sha_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for blk in iter(lambda: f.read(1 << 24), b""):
            h.update(blk)
    return h.hexdigest()


def _write_fvecs(path, X):
    X = np.ascontiguousarray(X, np.float32)
    rec = np.empty((X.shape[0], X.shape[1] + 1), np.float32)
    rec[:, 0] = np.array([X.shape[1]], np.int32).view(np.float32)[0]
    rec[:, 1:] = X
    rec.tofile(path)


def _write_seqs(path, atoms, offs):
    with open(path, "wb") as f:
        for i in range(offs.shape[0] - 1):
            f.write(atoms[offs[i]:offs[i + 1]].tobytes())
            f.write(b"\n")


def _run(driver, ranks, args, workdir):
    cmd = [os.path.join(MPI_HOME, "bin", "mpiexec.hydra"), "-n", str(ranks), os.path.join(REF, driver)] + args
    t = time.time()
    r = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True, timeout=7200)
    wall = time.time() - t
    if r.returncode != 0:
        raise RuntimeError(f"{driver} failed: {r.stderr[-2000:]}")
    timers = [ln for ln in (r.stderr + r.stdout).splitlines() if ln.startswith("[time=")]
    return wall, timers


def _graph_summary(path):
    """Header counts and the SHA-256 of the sorted upper-triangle pairs (i < j,
    int64) of a graph file in the reference's MPI-writer format."""
    import pandas as pd
    with open(path) as f:
        hdr = f.readline().split()
    df = pd.read_csv(path, sep=" ", skiprows=1, header=None, names=["u", "v", "w"],
                     dtype={"u": np.int64, "v": np.int64, "w": str})
    u, v = df["u"].to_numpy(), df["v"].to_numpy()
    up = u < v
    ij = np.stack([u[up], v[up]], 1)
    ij = ij[np.lexsort((ij[:, 1], ij[:, 0]))]
    return {"header": " ".join(hdr), "lines": int(u.shape[0]), "edges_undirected": int(ij.shape[0]),
            "sha256_pairs_int64": hashlib.sha256(np.ascontiguousarray(ij, np.int64).tobytes()).hexdigest()}


def _entry(driver, ranks, flags, inp, out, wall, timers, extra=None):
    e = {"driver": driver, "ranks": ranks, "flags": flags, "input_sha256": _sha_file(inp),
         "output_sha256": _sha_file(out), "output_bytes": os.path.getsize(out), "wall_s": round(wall, 2),
         "timers": timers}
    if extra:
        e.update(extra)
    return e


def job_c1(res, wd):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_small()
    inp = os.path.join(wd, "gmm_small.fvecs")
    _write_fvecs(inp, pts)
    r = "%.9g" % float(eps)
    for driver, flags in (("systolic", ["-r", r]), ("landmark", ["-r", r, "-m", "25", "-s", "7"])):
        out = os.path.join(wd, f"c1_{driver}.txt")
        wall, timers = _run(driver, 8, ["-i", inp, "-o", out] + flags, wd)
        res[f"{driver}_gmm_small_10kx16"] = _entry(driver, 8, flags, inp, out, wall, timers, _graph_summary(out))


def job_edit600(res, wd):
    from rnng_amd import synth
    atoms, offs = synth.dna_mutants(n_seeds=6, per_seed=100)
    inp = os.path.join(wd, "dna600.seqs")
    _write_seqs(inp, atoms, offs)
    # landmark: 3 x 8 ranks = 24 centers, two of them equal strings (an empty
    # cell), and 3 centers (-F), all distinct
    for key, driver, flags in (("systolic_edit_600", "systolic", ["-r", "8", "-D", "edit"]),
                               ("landmark_edit_600", "landmark", ["-r", "8", "-D", "edit", "-m", "3", "-s", "5"]),
                               ("landmark_edit_600_m3", "landmark",
                                ["-r", "8", "-D", "edit", "-m", "3", "-F", "-s", "5"])):
        out = os.path.join(wd, f"{key}.txt")
        wall, timers = _run(driver, 8, ["-i", inp, "-o", out] + flags, wd)
        res[key] = _entry(driver, 8, flags, inp, out, wall, timers, _graph_summary(out))


def job_edit40k(res, wd):
    from rnng_amd import synth
    atoms, offs = synth.dna_mutants()
    n = 40_000
    inp = os.path.join(wd, "dna40k.seqs")
    _write_seqs(inp, atoms[: int(offs[n])], offs[: n + 1])
    flags = ["-r", "8", "-D", "edit"]
    out = os.path.join(wd, "edit40k_systolic.txt")
    wall, timers = _run("systolic", 8, ["-i", inp, "-o", out] + flags, wd)
    res["systolic_edit_40k"] = _entry("systolic", 8, flags, inp, out, wall, timers, _graph_summary(out))


def job_rips(res, wd):
    from rnng_amd import synth
    for n, eps, dim, vform in ((5000, 0.3, 2, False), (1500, 0.35, 3, True)):
        pts, _ = synth.points_rips(n=n)
        inp = os.path.join(wd, f"rips{n}.fvecs")
        _write_fvecs(inp, pts)
        flags = ["-r", "%.9g" % float(np.float32(eps)), "-d", str(dim), "-m", "4", "-s", "11"] + (["-V"] if vform else [])
        base = os.path.join(wd, f"rips{n}")
        wall, timers = _run("rips", 8, ["-i", inp, "-o", base] + flags, wd)
        out = base + ".txt"
        with open(out) as f:
            nlines = sum(1 for _ in f)
        res[f"rips_{n}_d{dim}"] = _entry("rips", 8, flags, inp, out, wall, timers, {"simplices": nlines})


def job_gmm_int(res, wd):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_int()
    inp = os.path.join(wd, "gmm_int.fvecs")
    _write_fvecs(inp, pts)
    del pts
    flags = ["-r", "%.9g" % float(eps)]
    out = os.path.join(wd, "gmm_int_systolic.txt")
    wall, timers = _run("systolic", 8, ["-i", inp, "-o", out] + flags, wd)
    res["systolic_gmm_int_1Mx128"] = _entry("systolic", 8, flags, inp, out, wall, timers, _graph_summary(out))


def job_c2scale(res, wd):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_int()
    flags = ["-r", "%.9g" % float(eps)]
    for n in (50_000, 100_000, 200_000):
        inp = os.path.join(wd, f"gmm_int_{n}.fvecs")
        _write_fvecs(inp, pts[:n])
        out = os.path.join(wd, f"gmm_int_{n}.txt")
        wall, timers = _run("systolic", 8, ["-i", inp, "-o", out] + flags, wd)
        res[f"systolic_gmm_int_{n // 1000}k"] = _entry("systolic", 8, flags, inp, out, wall, timers,
                                                       _graph_summary(out))
        os.remove(out)


def job_c4sub(res, wd):
    from rnng_amd import synth
    pts, eps = synth.points_vmf()
    inp = os.path.join(wd, "vmf200k.fvecs")
    _write_fvecs(inp, pts[:200_000])
    del pts
    flags = ["-r", "%.9g" % float(eps), "-m", "410", "-F", "-s", "7"]
    out = os.path.join(wd, "vmf200k_landmark.txt")
    wall, timers = _run("landmark", 8, ["-i", inp, "-o", out] + flags, wd)
    res["landmark_vmf_200kx96"] = _entry("landmark", 8, flags, inp, out, wall, timers, _graph_summary(out))
    os.remove(out)


def job_c5cells(res, wd):
    from rnng_amd import synth
    pts, eps = synth.points_rips()
    inp = os.path.join(wd, "rips300k.fvecs")
    _write_fvecs(inp, pts)
    flags = ["-r", "%.9g" % float(np.float32(eps)), "-d", "2", "-m", "32", "-s", "11"]
    base = os.path.join(wd, "rips300k")
    wall, timers = _run("rips", 8, ["-i", inp, "-o", base] + flags, wd)
    out = base + ".txt"
    with open(out) as f:
        nlines = sum(1 for _ in f)
    res["rips_300k_d2_256cells"] = _entry("rips", 8, flags, inp, out, wall, timers, {"simplices": nlines})
    for fn in os.listdir(wd):
        if fn.startswith("rips300k"):
            os.remove(os.path.join(wd, fn))


JOBS = {"c1": job_c1, "edit600": job_edit600, "rips": job_rips, "edit40k": job_edit40k, "gmm_int": job_gmm_int,
        "c2scale": job_c2scale, "c4sub": job_c4sub, "c5cells": job_c5cells}


def main(names):
    for b in ("systolic", "landmark", "rips"):
        if not os.path.exists(os.path.join(REF, b)):
            sys.exit(f"{os.path.join(REF, b)} is missing: build it with make -C oracle ref")
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as wd:
        for name in names:
            t = time.time()
            JOBS[name](res, wd)
            with open(OUT, "w") as f:
                json.dump(res, f, indent=1, sort_keys=True)
            print(f"{name}: done in {time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "edit600", "rips"])
