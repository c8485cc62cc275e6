"""Host-side checks of the C++ drop-in headers (include/): the device cache's validation logic,
compiled against the headers and the library without touching a GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "sparseops_amd")


def _build_and_run(src, tmp_path):
    exe = os.path.join(str(tmp_path), os.path.splitext(os.path.basename(src))[0])
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-I" + INC, "-o", exe, src, "-L" + LIBDIR,
                           "-lsparseops_hip", "-Wl,-rpath," + LIBDIR])
    return subprocess.run([exe], capture_output=True, text=True, timeout=60)


def test_dropin_cache_detects_changed_matrices(tmp_path):
    """An entry is reused only while dims, array addresses and the content digest match: a matrix
    freed and reallocated at the same address with other contents is rebuilt, an in-place edit
    is caught where the digest covers it (always in full mode), and off mode never caches."""
    r = _build_and_run(os.path.join(ROOT, "tests", "cpp", "dropin_cache.cpp"), tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin cache: ok" in r.stdout
