"""The reference's own command-line callers compile and link, unchanged, against the drop-in
include/ and libsparseops_hip.so: benchmark_spmv_{csr,ell,sell_c_sigma,sell_c_R,bsr}.cpp,
test_spgemm_hc_lsh.cpp, test_spgemm_lsh.cpp, record_permutation_lsh.cpp, test_bsr_format.cpp,
generate_candidate_pairs.cpp, test_spgemm_hc.cpp and Get_features.cpp from the reference's test/
directory.  Each source is compiled where it
lies, reached through a symlink in a scratch directory whose include/ links to this
repository's headers, so its "../include/X.h" lines resolve to the drop-in (the recipe of
oracle/Makefile's `callers` target).  Runs only where the reference sources are present."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TEST = "/root/reference/test"
LIB = os.path.join(ROOT, "sparseops_amd", "libsparseops_hip.so")
CALLERS = ["benchmark_spmv_csr", "benchmark_spmv_ell", "benchmark_spmv_sell_c_sigma", "benchmark_spmv_sell_c_R",
           "benchmark_spmv_bsr", "test_spgemm_hc_lsh", "test_spgemm_lsh", "record_permutation_lsh",
           "test_bsr_format", "generate_candidate_pairs", "test_spgemm_hc", "Get_features"]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_TEST), reason="the reference's callers are not present here")


def _compile(name, scratch):
    d = os.path.join(scratch, name)
    os.makedirs(os.path.join(d, "test"))
    os.symlink(os.path.join(REF_TEST, name + ".cpp"), os.path.join(d, "test", name + ".cpp"))
    os.symlink(os.path.join(ROOT, "include"), os.path.join(d, "include"))
    cmd = ["g++", "-std=c++17", "-O0", "-fopenmp", "-Wall", "-I/opt/rocm/include", os.path.join(d, "test", name + ".cpp"),
           "-o", os.path.join(d, name), "-L" + os.path.dirname(LIB), "-lsparseops_hip"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    return name, p.returncode, p.stderr, d


def test_reference_callers_compile_against_dropin(tmp_path):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda n: _compile(n, str(tmp_path)), CALLERS))
    failed = [(n, err[-2000:]) for n, rc, err, _ in results if rc != 0]
    assert not failed, failed
    # the drop-in headers themselves compile without warnings under -Wall
    warn = [(n, line) for n, _, err, d in results for line in err.splitlines()
            if "warning" in line and os.path.join(d, "test", "..", "include") in line]
    assert not warn, warn[:10]
    # every caller resolved its ../include/ lines to the drop-in headers, none to the reference's
    name, _, _, d = results[0]
    dep = subprocess.run(["g++", "-std=c++17", "-fopenmp", "-I/opt/rocm/include", "-M",
                          os.path.join(d, "test", name + ".cpp")], capture_output=True, text=True, timeout=300)
    assert dep.returncode == 0, dep.stderr
    assert "/root/reference/include" not in dep.stdout
    assert os.path.join(d, "test", "..", "include", "SpOps.h") in dep.stdout


def test_dropin_has_every_umbrella_header():
    """Every header the reference's umbrella SpOps.h pulls in for this library's formats
    exists in the drop-in include/ (CSR5, DIA, S-ELL, COO kernels and the accumulator
    micro-benchmarks are not part of this library)."""
    want = ["SpOps.h", "general_config.h", "plat_config.h", "memopt.h", "mmio.h", "thread.h", "timer.h",
            "cmdline.h", "sparse_io.h", "sparse_format.h", "sparse_operation.h", "sparse_partition.h",
            "sparse_conversion.h", "spmv_benchmark.h", "spmv_testroutine.h", "spgemm.h", "spgemm_utility.h",
            "spgemm_MinHashLSH.h", "spgemm_cluster.h", "sparse_features.h"]
    missing = [h for h in want if not os.path.exists(os.path.join(ROOT, "include", h))]
    assert not missing, missing
    assert shutil.which("g++")
