"""MAT-file (Level 5) I/O of the durable per-shot results, on the CPU: round
trips, files that scipy.io (an independent implementation of the published
format) reads, and files scipy.io writes -- uncompressed (save -v6) and
zlib-compressed (save -v7), with integer-valued data that a writer may store in
a narrower numeric type."""
import os

import numpy as np
import pytest
import scipy.io

import sssi_amd
from sssi_amd import mat_read, mat_write


def test_round_trip_several_variables(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 5))
    b = rng.standard_normal((3, 4, 6))
    c = np.arange(11.0)
    p = tmp_path / "x.mat"
    mat_write(p, dataTrue=a, snap=b, key=c)
    assert np.array_equal(mat_read(p, "dataTrue"), a)
    assert np.array_equal(mat_read(p, "snap"), b)
    assert np.array_equal(mat_read(p, "key").ravel(), c)
    assert [f for f in os.listdir(tmp_path) if "tmp" in f] == []  # written atomically


def test_scipy_reads_our_files(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((9, 13))
    b = rng.standard_normal((2, 3, 4))
    p = tmp_path / "y.mat"
    mat_write(p, image=a, stack=b)
    m = scipy.io.loadmat(str(p))
    assert np.array_equal(m["image"], a) and np.array_equal(m["stack"], b)


@pytest.mark.parametrize("compress", [False, True])
def test_we_read_scipy_files(tmp_path, compress):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 8))
    ints = np.arange(24.0).reshape(4, 6)  # integer-valued doubles
    p = tmp_path / "z.mat"
    scipy.io.savemat(str(p), {"StackedRTM": a, "n": ints, "s": 3.5}, do_compression=compress)
    assert np.array_equal(mat_read(p, "StackedRTM"), a)
    assert np.array_equal(mat_read(p, "n"), ints)
    assert mat_read(p, "s").ravel().tolist() == [3.5]


def test_missing_file_and_variable_are_errors(tmp_path):
    with pytest.raises(sssi_amd.SssiError) as e:
        mat_read(tmp_path / "none.mat", "x")
    assert e.value.code == -1
    p = tmp_path / "w.mat"
    mat_write(p, a=np.ones((2, 2)))
    with pytest.raises(sssi_amd.SssiError) as e:
        mat_read(p, "b")
    assert "no real double variable b" in str(e.value)
    q = tmp_path / "h5.mat"
    q.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(200))   # what -v7.3 (HDF5) files start with
    with pytest.raises(sssi_amd.SssiError) as e:
        mat_read(q, "x")
    assert "Level 5" in str(e.value)
