"""The reference's own command-line callers, compiled unchanged against the drop-in include/
(oracle/_ref/callers, built by oracle/Makefile's `callers` target where the reference sources
are present), run on the GPU as child processes:

  * benchmark_spmv_{csr,ell,sell_c_sigma,sell_c_R,bsr} --precision=64 and 32 on a power-law
    matrix: the test routine's check of SpOps_<format> (GPU) against a host CSR product prints
    "max error" without the reference's warnings, and the performance line is appended to
    MAT_PERFORMANCE in the reference's layout;
  * test_spgemm_hc_lsh --test_type=correctness (kernels 1 and 3): the C = A * A it writes in the
    original row order equals scipy's product (structure exact, values 1e-12 relative);
  * record_permutation_lsh then test_spgemm_lsh: the clustering written equals the reference's
    own (tests/golden) and the replay reports the reference's flop count.
"""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from sparseops_amd.gen import powerlaw, rmat
from sparseops_amd.io import read_clustering, read_mtx, write_mtx

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "callers")
GOLD = os.path.join(ROOT, "tests", "golden")


def _run(cwd, name, *args, timeout=300):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} is not built: the reference's callers are compiled where its sources are present")
    print(f"run {name} {' '.join(args)}", flush=True)
    os.makedirs(os.path.join(cwd, "performance"), exist_ok=True)
    os.makedirs(os.path.join(cwd, "features"), exist_ok=True)
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)


@pytest.fixture(scope="module")
def pl_mtx(tmp_path_factory):
    d = tmp_path_factory.mktemp("pl")
    A = powerlaw(3000, seed=3, lmax=300)
    path = str(d / "pl3000.mtx")
    write_mtx(path, A)
    return path


@pytest.mark.parametrize("precision", ["64", "32"])
@pytest.mark.parametrize("fmt,tag", [("csr", "CSR"), ("ell", "ELL"), ("sell_c_sigma", "S-ELL-sigma"),
                                     ("sell_c_R", "S-ELL-R"), ("bsr", "BSR")])
def test_benchmark_spmv_callers(pl_mtx, fmt, tag, precision, tmp_path):
    p = _run(str(tmp_path), f"benchmark_spmv_{fmt}", pl_mtx, f"--precision={precision}", "--matID=7")
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    errs = [float(e) for e in re.findall(r"\[max error\s+([0-9.eE+-]+)\]", p.stdout)]
    assert errs, p.stdout[-2000:]
    assert max(errs) <= (1e-6 if precision == "64" else 1e-4), errs
    assert "POSSIBLE" not in p.stdout
    assert re.search(r"benchmarking \S+\s+\[gpu\]: +[0-9.]+ ms \( [0-9.]+ GFLOP/s [0-9.]+ GB/s\)", p.stdout)
    perf = open(tmp_path / "performance" / "mat_perf.txt").read().split()
    assert perf[:3] == ["7", "pl3000", tag], perf


@pytest.mark.parametrize("kernel", ["1", "3"])
def test_spgemm_hc_lsh_caller_correctness(kernel, tmp_path):
    A = rmat(10, 16, seed=10)
    mtx = str(tmp_path / "rmat10.mtx")
    write_mtx(mtx, A)
    p = _run(str(tmp_path), "test_spgemm_hc_lsh", mtx, mtx, "--test_type=correctness", f"--kernel={kernel}",
             "--precision=64")
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "Correctness test done." in p.stdout
    suffix = "hashvlengthcluster_hc_lsh" if kernel == "1" else "mixedvlengthcluster_hc_lsh"
    C = read_mtx(str(tmp_path / f"rmat10_SpOps_{suffix}.mtx"))
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
    want = (S @ S).tocsr()
    want.sort_indices()
    got = sp.csr_matrix((C.val, C.col, C.rowptr), shape=(C.n_rows, C.n_cols))
    got.sum_duplicates()
    got.sort_indices()
    np.testing.assert_array_equal(got.indptr, want.indptr)
    np.testing.assert_array_equal(got.indices, want.indices)
    np.testing.assert_allclose(got.data, want.data, rtol=1e-12, atol=0)


def test_record_permutation_then_spgemm_lsh_callers(tmp_path):
    A = rmat(10, 16, seed=10)
    mtx = str(tmp_path / "rmat10.mtx")
    write_mtx(mtx, A)
    out = str(tmp_path / "lsh")
    p = _run(str(tmp_path), "record_permutation_lsh", mtx, f"--out-dir={out}")
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    g = np.load(os.path.join(GOLD, "lsh_rmat10.npz"))
    perm, offsets = read_clustering(out, "rmat10", A.n_rows)
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(offsets, g["offsets"])
    p = _run(str(tmp_path), "test_spgemm_lsh", mtx, mtx, f"--lsh-order-dir={out}", "--iterations=2")
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert int(re.search(r"FLOPs: (\d+)", p.stdout).group(1)) == int(g["flops"][0])
    assert int(re.search(r"Clusters: (\d+)", p.stdout).group(1)) == len(offsets) - 1
