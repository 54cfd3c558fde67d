"""CPU checks of the C++ host layer (include/rnng/*.h) and the command-line
drivers: headers compile warning-free, the drivers take the reference's
flags, simplex ids / vertex lists match the combinatorial number system, and
the graph and point-file formats match the reference's writers and readers.
No GPU call."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "rnng_amd", "bin")


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("host") / "host_checks")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host", "host_checks.cpp"), "-L", os.path.join(ROOT, "rnng_amd"),
                    "-lrnng_amd", "-Wl,-rpath," + os.path.join(ROOT, "rnng_amd"), "-o", out], check=True)
    return out


@pytest.mark.parametrize("prog,flags", [("rnng_systolic", "i:r:c:l:v:o:D:h"),
                                        ("rnng_landmark", "i:r:m:c:l:v:o:s:FD:h"),
                                        ("rnng_rips", "i:r:m:c:l:v:o:s:FD:d:Vh")])
def test_drivers_print_usage(prog, flags):
    path = os.path.join(BIN, prog)
    assert os.path.exists(path), "build the drivers first (make -C rnng_amd/csrc)"
    r = subprocess.run([path, "-h"], capture_output=True, text=True)
    assert r.returncode == 0
    assert "-i <points> -r <radius>" in r.stderr
    for opt in flags.replace(":", ""):
        if opt not in "ir":
            assert f"-{opt} " in r.stderr, opt
    # missing input is a usage error
    assert subprocess.run([path], capture_output=True).returncode == 1


def test_simplex_ids_and_repr(host, orc):
    rng = np.random.default_rng(0)
    cases = [[0], [5], [0, 1], [3, 9], [7, 2, 11], [999999, 3, 500000]]
    cases += [sorted(rng.choice(1_000_000, size=k, replace=False).tolist()) for k in (1, 2, 3) for _ in range(20)]
    inp = "".join(f"{len(c)} " + " ".join(map(str, c)) + "\n" for c in cases)
    out = subprocess.run([host, "simplex"], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        sid, dim, rep = line.split()
        assert int(sid) == orc.simplex_id(np.array(c))
        assert int(dim) == len(c) - 1
        assert rep == "<" + ",".join(map(str, sorted(c))) + ">"


def test_graph_writers(host, tmp_path):
    edges = [(0, 3, 1.5), (3, 0, 1.5), (1, 2, 0.123456789), (2, 2, 0.0), (4, 1, 1e-7), (0, 1, 256.289)]
    inp = f"5 {len(edges)}\n" + "".join(f"{u} {v} {w!r}\n" for u, v, w in edges)
    a, b = tmp_path / "a.txt", tmp_path / "b.txt"
    subprocess.run([host, "graph", str(a), str(b)], input=inp, text=True, check=True)
    # both orientations, sorted, unique (include/graph.hpp:59-116)
    sym = sorted({(u, v, float(np.float32(w))) for u, v, w in edges} |
                 {(v, u, float(np.float32(w))) for u, v, w in edges})
    want_a = [f"% 5 5 {len(sym)}"] + ["%d %d %.6f" % e for e in sym]
    want_b = [f"% 5 5 {len(sym)}"] + ["%d %d %s" % (u, v, "%.6g" % w) for u, v, w in sym]
    assert a.read_text().splitlines() == want_a
    assert b.read_text().splitlines() == want_b


def test_fvecs_reader(host, tmp_path):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((7, 5)).astype(np.float32)
    p = tmp_path / "x.fvecs"
    with open(p, "wb") as f:
        for row in X:
            f.write(struct.pack("<i", 5))
            f.write(row.tobytes())
    out = subprocess.run([host, "fvecs", str(p)], capture_output=True, text=True, check=True).stdout.splitlines()
    assert out[0] == "7 5"
    for i, line in enumerate(out[1:]):
        vals = line.split()
        assert int(vals[0]) == i
        assert np.allclose(np.array(vals[1:], dtype=np.float64), X[i], rtol=1e-5)


@pytest.mark.parametrize("ext", [".fbin", ".u8bin", ".fvecs"])
def test_read_file_by_extension(host, tmp_path, ext):
    """PointContainer::read_file: .fbin and .u8bin headers (uint32 n, d) and
    bodies, u8 values as floats, anything else as fvecs."""
    from rnng_amd import dataset_io
    rng = np.random.default_rng(2)
    X = rng.integers(0, 256, size=(9, 4))
    if ext != ".u8bin":
        X = X / 8.0   # exactly representable, printed exactly
    p = str(tmp_path / ("x" + ext))
    dataset_io.write_file(p, X.astype(np.uint8 if ext == ".u8bin" else np.float32))
    out = subprocess.run([host, "file", p], capture_output=True, text=True, check=True).stdout.splitlines()
    assert out[0] == "9 4"
    for i, line in enumerate(out[1:]):
        vals = line.split()
        assert int(vals[0]) == i
        assert np.array_equal(np.array(vals[1:], dtype=np.float64), X[i])


def test_seqs_reader(host, tmp_path):
    p = tmp_path / "x.seqs"
    p.write_bytes(b"ACGT\r\n\nGATTACA\n\r\nTT\n")
    out = subprocess.run([host, "seqs", str(p)], capture_output=True, text=True, check=True).stdout.splitlines()
    # empty lines (also a lone '\r') are skipped, '\r' is dropped
    assert out == ["3", "0 ACGT", "1 GATTACA", "2 TT"]
