"""Point-file readers and writers of the reference's formats (dataset_io.py:
u8bin, fbin, fvecs), with start/count slicing, on the CPU."""
import numpy as np
import pytest

from rnng_amd import dataset_io as io


@pytest.mark.parametrize("ext,dtype", [(".u8bin", np.uint8), (".fbin", np.float32), (".fvecs", np.float32)])
def test_round_trip_and_slices(tmp_path, ext, dtype):
    rng = np.random.default_rng(1)
    X = (rng.integers(0, 256, size=(53, 7)) if dtype == np.uint8 else rng.standard_normal((53, 7))).astype(dtype)
    f = str(tmp_path / ("p" + ext))
    io.write_file(f, X)
    n, d, size, kind = io.info_file(f)
    assert (int(n), int(d)) == (53, 7)
    assert kind == ("uint8" if dtype == np.uint8 else "float32")
    assert size == {".u8bin": 8 + 53 * 7, ".fbin": 8 + 53 * 7 * 4, ".fvecs": 53 * 8 * 4}[ext]
    assert np.array_equal(io.read_file(f), X)
    assert np.array_equal(io.read_file(f, 10), X[10:])
    assert np.array_equal(io.read_file(f, 10, 5), X[10:15])
    assert np.array_equal(io.read_file(f, 0, 1), X[:1])


def test_fvecs_layout_is_dimension_then_floats(tmp_path):
    X = np.arange(6, dtype=np.float32).reshape(2, 3)
    f = str(tmp_path / "p.fvecs")
    io.write_fvecs(f, X)
    raw = np.fromfile(f, dtype=np.int32).reshape(2, 4)
    assert list(raw[:, 0]) == [3, 3]
    assert np.array_equal(raw[:, 1:].view(np.float32), X)


def test_unknown_extension(tmp_path):
    with pytest.raises(ValueError):
        io.read_file(str(tmp_path / "p.txt"))
    with pytest.raises(ValueError):
        io.write_file(str(tmp_path / "p.npy"), np.zeros((1, 1), np.float32))
