"""The reference's command-line drivers on the drop-in headers (sparseops_amd/cli), run as
child processes on the GPU: record_permutation_lsh writes the clustering of
test/record_permutation_lsh.cpp (compared with the reference's own permutation and offsets
in tests/golden), test_spgemm_lsh reloads it and times LeSpGEMM_VLength kernel 3
(test/test_spgemm_lsh.cpp), printing the reference's stdout lines."""
import os
import re
import subprocess

import numpy as np
import pytest

from sparseops_amd.gen import rmat
from sparseops_amd.io import read_clustering, write_clustering, write_mtx

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sparseops_amd", "cli")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(*args):
    exe = os.path.join(CLI, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C sparseops_amd/cli)")
    return subprocess.run([exe, *args[1:]], capture_output=True, text=True, timeout=300)


def test_record_then_replay_clustering(tmp_path):
    A = rmat(10, 16, seed=10)
    mtx = str(tmp_path / "rmat10.mtx")
    write_mtx(mtx, A)
    out = str(tmp_path / "lsh")
    p = _run("record_permutation_lsh", mtx, f"--out-dir={out}")
    assert p.returncode == 0, p.stderr
    assert "LSH candidate pairs:" in p.stdout and "Wrote" in p.stdout
    g = np.load(os.path.join(GOLD, "lsh_rmat10.npz"))
    perm, offsets = read_clustering(out, "rmat10", A.n_rows)
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(offsets, g["offsets"])
    p = _run("test_spgemm_lsh", mtx, mtx, f"--lsh-order-dir={out}", "--iterations=2")
    assert p.returncode == 0, p.stderr
    flops = int(re.search(r"FLOPs: (\d+)", p.stdout).group(1))
    assert flops == int(g["flops"][0])
    assert int(re.search(r"Clusters: (\d+)", p.stdout).group(1)) == len(offsets) - 1
    assert int(re.search(r"nnzc=(\d+)", p.stdout).group(1)) == int(g["acl_meta"][1])
    assert float(re.search(r"Average GFLOPS: ([0-9.e+-]+)", p.stdout).group(1)) > 0


def test_saved_clustering_round_trip_and_errors(tmp_path):
    A = rmat(8, 16, seed=8)
    mtx = str(tmp_path / "r8.mtx")
    write_mtx(mtx, A)
    p = _run("record_permutation_lsh", mtx, f"--out-dir={tmp_path}/none", "--save=0")
    assert p.returncode == 0 and "Save disabled" in p.stdout and not os.path.exists(f"{tmp_path}/none/r8.perm")
    # a clustering written from Python is what test_spgemm_lsh reads (identity permutation, singletons)
    write_clustering(str(tmp_path / "py"), "r8", np.arange(A.n_rows), np.arange(A.n_rows + 1))
    p = _run("test_spgemm_lsh", mtx, mtx, f"--lsh-order-dir={tmp_path}/py", "--iterations=1")
    assert p.returncode == 0, p.stderr
    assert int(re.search(r"Clusters: (\d+)", p.stdout).group(1)) == A.n_rows
    # offsets that do not end at n_rows are refused
    write_clustering(str(tmp_path / "bad"), "r8", np.arange(A.n_rows), np.array([0, 5]))
    p = _run("test_spgemm_lsh", mtx, mtx, f"--lsh-order-dir={tmp_path}/bad", "--iterations=1")
    assert p.returncode != 0 and "offsets must be non-empty" in p.stderr
