"""Point files straight into HBM (rnng_dataset_read_file): the device copy of
every format, slice and chunking equals the host reader's array bit for bit,
and malformed files or ranges are refused with the documented codes."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _io():
    from rnng_amd import dataset_io
    return dataset_io


@pytest.mark.parametrize("ext", [".u8bin", ".fbin", ".fvecs"])
@pytest.mark.parametrize("n,d", [(1, 1), (1000, 3), (300_000, 128)])
def test_device_load_equals_host_read(tmp_path, ext, n, d):
    io = _io()
    rng = np.random.default_rng(n + d)
    if ext == ".u8bin":
        X = rng.integers(0, 256, size=(n, d)).astype(np.uint8)
    else:
        X = rng.standard_normal((n, d)).astype(np.float32)
        X.reshape(-1)[:: max(1, X.size // 7)] = np.float32(-0.0)
    f = str(tmp_path / ("p" + ext))
    io.write_file(f, X)
    flush_print(f"ingest {ext} {n}x{d}")
    assert io.device_info(f) == (n, d, 1 if ext == ".u8bin" else 0)
    for start, count in [(0, None), (n // 3, None), (n // 2, max(1, n // 5)), (n - 1, 1)]:
        ds = io.load_device(f, start, count)
        got = ds.to_numpy()
        want = io.read_file(f, start, count).astype(np.float32)
        assert got.shape == want.shape
        assert got.tobytes() == want.tobytes()
        ds.close()


def test_device_load_graph_equals_array_graph(tmp_path):
    from rnng_amd import l2, synth
    io = _io()
    pts, eps = synth.points_gmm_small(n=4000)
    f = str(tmp_path / "p.fbin")
    io.write_fbin(f, pts)
    a = io.load_device(f)
    b = l2.Dataset(pts)
    ij_a, w_a = a.graph(eps, "mfma").edges()
    ij_b, w_b = b.graph(eps, "mfma").edges()
    assert np.array_equal(ij_a, ij_b) and w_a.tobytes() == w_b.tobytes()
    assert ij_a.shape[0] > 1000


def test_device_load_errors(tmp_path):
    from rnng_amd._lib import RNNG_ERR_ARG, RNNG_ERR_FORMAT, RnngError
    io = _io()
    X = np.ones((10, 4), np.float32)
    f = str(tmp_path / "p.fbin")
    io.write_fbin(f, X)
    for args, code in [((f, 11, None), RNNG_ERR_ARG), ((f, 5, 6), RNNG_ERR_ARG),
                       ((str(tmp_path / "p.txt"), 0, None), RNNG_ERR_FORMAT)]:
        if args[0].endswith(".txt"):
            open(args[0], "w").write("x")
        with pytest.raises(RnngError) as ei:
            io.load_device(*args)
        assert ei.value.code == code
    # truncated body
    t = str(tmp_path / "t.fbin")
    with open(t, "wb") as fh:
        np.array([10, 4], np.uint32).tofile(fh)
        np.ones(39, np.float32).tofile(fh)
    with pytest.raises(RnngError) as ei:
        io.load_device(t)
    assert ei.value.code == RNNG_ERR_FORMAT
    # fvecs record with another dimension
    v = str(tmp_path / "v.fvecs")
    io.write_fvecs(v, X)
    raw = np.fromfile(v, dtype=np.int32)
    raw[5 * 5] = 3
    raw.tofile(v)
    with pytest.raises(RnngError) as ei:
        io.load_device(v)
    assert ei.value.code == RNNG_ERR_FORMAT


def _seqs_reference(data: bytes):
    """The reference's reader (include/point.hpp:170-205) restated: getline on
    '\\n', one trailing '\\r' dropped, empty lines skipped."""
    parts = data.split(b"\n")
    if parts and parts[-1] == b"":
        parts.pop()
    out = []
    for ln in parts:
        if ln.endswith(b"\r"):
            ln = ln[:-1]
        if ln:
            out.append(ln)
    return out


def _csr(strings):
    offs = np.zeros(len(strings) + 1, np.int64)
    offs[1:] = np.cumsum([len(s) for s in strings])
    return np.frombuffer(b"".join(strings), np.uint8), offs


@pytest.mark.parametrize("case", ["edges", "dna", "big"])
def test_seqs_device_read_equals_reference_reader(tmp_path, case):
    from rnng_amd import edit, synth
    rng = np.random.default_rng(11)
    if case == "edges":
        data = (b"ACGT\r\n\r\n\n\rAC\r\r\nTTTT\n\r\n" + bytes(range(1, 256)).replace(b"\n", b"") + b"\nlast-no-newline")
    elif case == "dna":
        atoms, offs = synth.dna_mutants(n_seeds=300, per_seed=20)
        lines = [atoms[offs[i]:offs[i + 1]].tobytes() for i in range(offs.shape[0] - 1)]
        data = b"".join(l + (b"\r\n" if i % 3 == 0 else b"\n") + (b"\n" if i % 7 == 0 else b"")
                        for i, l in enumerate(lines))
    else:
        # above two 32 MiB pinned chunks: the double-buffered upload wraps
        lines = [bytes(rng.choice(list(b"ACGT"), int(L))) for L in rng.integers(100, 151, 600_000)]
        data = b"\n".join(lines) + b"\n"
    f = tmp_path / "s.seqs"
    f.write_bytes(data)
    want = _seqs_reference(data)
    flush_print(f"seqs {case}: {len(data)} bytes, {len(want)} strings")
    for start, count in [(0, None), (len(want) // 3, None), (1, max(1, len(want) // 4)), (len(want), 0)]:
        S = edit.StringSet.from_file(str(f), start, count)
        sub = want[start:] if count is None else want[start:start + count]
        a, o = _csr(sub)
        assert S.n == len(sub)
        assert np.array_equal(S.offsets, o) and S.atoms.tobytes() == a.tobytes()
        S.close()
    from rnng_amd import _lib
    import ctypes as C
    h = C.c_void_p()
    assert _lib.lib().rnng_strings_read_file(str(f).encode(), len(want) + 1, -1, C.byref(h)) == _lib.RNNG_ERR_ARG


def test_seqs_device_read_graph_equals_host_created(tmp_path):
    from rnng_amd import edit, synth
    atoms, offs = synth.dna_mutants(n_seeds=40, per_seed=50)
    f = tmp_path / "s.seqs"
    f.write_bytes(b"".join(atoms[offs[i]:offs[i + 1]].tobytes() + b"\n" for i in range(offs.shape[0] - 1)))
    A = edit.StringSet.from_file(str(f))
    B = edit.StringSet(atoms, offs)
    ga, gb = A.graph(8.0), B.graph(8.0)
    ija, wa = ga.edges()
    ijb, wb = gb.edges()
    assert ija.shape[0] > 1000 and np.array_equal(ija, ijb) and wa.tobytes() == wb.tobytes()
    for x in (ga, gb, A, B):
        x.close()
