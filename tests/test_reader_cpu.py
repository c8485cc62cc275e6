"""The MatrixMarket readers against the reference's read_csr_matrix (utils/sparse_io.cpp:26-214)
on the files of tests/golden/mtx/: the drop-in C++ read_csr_matrix (include/sparse_io.h) and the
Python read_mtx (sparseops_amd/io.py) give the reference's row pointer, columns and values for
compact = false (entries in file order inside a row, duplicates kept, explicit zeros kept) and
compact = true (duplicates summed, exact-zero sums dropped, columns in the reference's reversed
first-appearance order), with symmetric files mirrored (diagonal once), pattern entries 1.0,
integer fields read as values, skew-symmetric files left unmirrored (the reference mirrors only
"symmetric") and the banner matched case-insensitively.  tests/golden/reader_mtx.npz holds the
reference's output (tests/golden/make_golden.py reader)."""
import os
import subprocess

import numpy as np
import pytest

from sparseops_amd.io import read_mtx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MTX = os.path.join(ROOT, "tests", "golden", "mtx")
G = np.load(os.path.join(ROOT, "tests", "golden", "reader_mtx.npz"), allow_pickle=False)
FILES = sorted(f[:-4] for f in os.listdir(MTX) if f.endswith(".mtx"))


@pytest.fixture(scope="module")
def reader_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("reader")
    exe = os.path.join(str(d), "read_mtx_dump")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "read_mtx_dump.cpp"),
                           "-L" + os.path.join(ROOT, "sparseops_amd"), "-lsparseops_hip",
                           "-Wl,-rpath," + os.path.join(ROOT, "sparseops_amd")])
    return exe


@pytest.mark.parametrize("compact", [0, 1])
@pytest.mark.parametrize("name", FILES)
def test_dropin_reader_matches_reference(reader_exe, name, compact):
    r = subprocess.run([reader_exe, os.path.join(MTX, name + ".mtx"), str(compact)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip("\n").split("\n")
    dims = np.array(lines[0].split(), np.int64)
    np.testing.assert_array_equal(dims, G[f"{name}_c{compact}_dims"])
    np.testing.assert_array_equal(np.array(lines[1].split(), np.int64), G[f"{name}_c{compact}_rowptr"])
    np.testing.assert_array_equal(np.array(lines[2].split(), np.int64), G[f"{name}_c{compact}_col"])
    np.testing.assert_array_equal(np.array(lines[3].split(), np.float64), G[f"{name}_c{compact}_val"])


@pytest.mark.parametrize("compact", [0, 1])
@pytest.mark.parametrize("name", FILES)
def test_python_reader_matches_reference(name, compact):
    A = read_mtx(os.path.join(MTX, name + ".mtx"), compact=bool(compact))
    np.testing.assert_array_equal([A.n_rows, A.n_cols, A.nnz], G[f"{name}_c{compact}_dims"])
    np.testing.assert_array_equal(A.rowptr, G[f"{name}_c{compact}_rowptr"])
    np.testing.assert_array_equal(A.col, G[f"{name}_c{compact}_col"])
    np.testing.assert_array_equal(A.val, G[f"{name}_c{compact}_val"])
