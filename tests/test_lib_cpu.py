"""CPU-side checks of the C-ABI library: it loads, exports every symbol include/*.h declares,
maps bad arguments to status codes, and its host-only entry points (hierarchical
clustering) match the reference fixtures.  No GPU compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_bind as orc
from sparseops_amd import _lib as L
from sparseops_amd.gen import Csr, rmat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_in_headers():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h == "spops_hip.h":
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spops_\w+)\s*\(", txt, re.M))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_in_headers()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(L.lib, n)]
    assert not missing, missing
    assert set(L.declared_symbols()) == names, "ctypes signature table out of sync with include/*.h"


def test_status_strings_and_version():
    assert L.lib.spops_status_string(0) == b"success"
    assert L.lib.spops_status_string(5) == b"size limit exceeded"
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.lib.spops_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0


def test_invalid_arguments_map_to_status_codes():
    h = ctypes.c_void_p()
    rp = np.array([0, 1, 2], np.int64)
    ci = np.array([0, 5], np.int64)      # column 5 out of range for n_cols = 2
    v = np.ones(2)
    st = L.lib.spops_csr_create(ctypes.byref(h), 2, 2, 2, L.INDEX_I64, rp.ctypes.data, ci.ctypes.data,
                                v.ctypes.data, L.VALUE_F64, L.VALUE_F64)
    assert st == 1 and b"out of range" in L.lib.spops_last_error()
    st = L.lib.spops_sell_create(ctypes.byref(h), 2, 2, 2, L.INDEX_I64, rp.ctypes.data, rp.ctypes.data,
                                 v.ctypes.data, L.VALUE_F64, L.VALUE_F64, 10, 4)
    assert st == 1 and b"multiple" in L.lib.spops_last_error()
    st = L.lib.spops_csr_create(ctypes.byref(h), 2, 2, 2, 7, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                L.VALUE_F64, L.VALUE_F64)
    assert st == 1
    st = L.lib.spops_bsr_create(ctypes.byref(h), 2, 2, 2, L.INDEX_I64, rp.ctypes.data, rp.ctypes.data,
                                v.ctypes.data, L.VALUE_F64, L.VALUE_F64, 80, 80)
    assert st == 5                       # neither the VALU (<= 256 entries) nor the MFMA kernel (<= 64 rows)
    assert L.lib.spops_spmv(None, 1.0, None, 0.0, None, None) == 1
    assert L.lib.spops_hierarchical_clustering(3, 0, None, None, 0, None, None, None, 0, 8, None, None, None) == 1


@pytest.mark.parametrize("name,version,f32", [("lsh_rmat10", 0, False), ("lsh_rmat10_hc1", 1, False),
                                              ("lsh_rmat10_hc2", 0, False), ("lsh_rmat10_f32", 0, True),
                                              ("lsh_rmat12", 0, False)])
def test_library_hc_matches_reference(name, version, f32):
    from sparseops_amd.lespgemm import hierarchical_clustering
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    scale = 12 if name == "lsh_rmat12" else 10
    A = rmat(scale, 16, seed=scale)
    sig = orc.minhash(A, int(g["k"]), int(g["seed"]))
    pi, pj, ps = orc.lsh_pairs(sig, int(g["k"]), int(g["bands"]), constrain=int(g["hc_v"]) == 0)
    perm, off = hierarchical_clustering(A, pi, pj, ps.astype(np.float32) if f32 else ps, int(g["cluster_size"]),
                                        version)
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(off, g["offsets"])


def test_library_hc_matches_oracle_on_other_inputs():
    from sparseops_amd.lespgemm import hierarchical_clustering
    for scale, cs in [(11, 4), (13, 8), (12, 16)]:
        A = rmat(scale, 8, seed=100 + scale)
        sig = orc.minhash(A, 32, 3)
        pi, pj, ps = orc.lsh_pairs(sig, 32, 8)
        for version in (0, 1):
            want = orc.hc(A, pi, pj, ps, cs, version=version)
            got = hierarchical_clustering(A, pi, pj, ps, cs, version)
            np.testing.assert_array_equal(got[0], want[0])
            np.testing.assert_array_equal(got[1], want[1])


def test_library_hc_fallbacks_match_oracle():
    """Pairs in shuffled order (the search for a discovered pair among the input runs over
    all of them, as the reference's lower_bound does) and rows out of order or with repeated
    columns (the Jaccard of such rows takes set semantics): same clustering as the oracle."""
    from sparseops_amd.lespgemm import hierarchical_clustering
    A = rmat(11, 8, seed=7)
    sig = orc.minhash(A, 32, 3)
    pi, pj, ps = orc.lsh_pairs(sig, 32, 8)
    shuffled = np.random.default_rng(0).permutation(len(pi))
    rows = [A.col[A.rowptr[r]:A.rowptr[r + 1]].astype(np.int64) for r in range(A.n_rows)]
    for r in range(A.n_rows):
        if r % 7 == 0:
            rows[r] = rows[r][::-1].copy()
        if r % 11 == 0 and len(rows[r]):
            rows[r] = np.concatenate([rows[r], rows[r][:1]])
    rp = np.zeros(A.n_rows + 1, np.int64)
    np.cumsum([len(c) for c in rows], out=rp[1:])
    col = np.concatenate(rows)
    B = Csr(A.n_rows, A.n_cols, rp, col, np.ones(len(col)))
    for version in (0, 1):
        for M, qi, qj, qs in ((A, pi[shuffled], pj[shuffled], ps[shuffled]), (B, pi, pj, ps)):
            want = orc.hc(M, qi, qj, qs, 4, version=version)
            got = hierarchical_clustering(M, qi, qj, qs, 4, version)
            np.testing.assert_array_equal(got[0], want[0])
            np.testing.assert_array_equal(got[1], want[1])


def test_clustering_files_round_trip(tmp_path):
    """.perm / .offsets as record_permutation_lsh writes them; the reader pads a short
    permutation with (size, size+1, ...), cuts a long one, and skips a Matrix Market header."""
    from sparseops_amd.io import read_clustering, write_clustering
    perm = np.array([3, 0, 2, 1])
    pp, po = write_clustering(str(tmp_path), "m", perm, [0, 2, 4])
    assert open(pp).read() == "3\n0\n2\n1\n" and open(po).read() == "0\n2\n4\n"
    p, o = read_clustering(str(tmp_path), "m", 4)
    assert p.tolist() == [3, 0, 2, 1] and o.tolist() == [0, 2, 4]
    assert read_clustering(str(tmp_path), "m", 6)[0].tolist() == [3, 0, 2, 1, 4, 5]
    assert read_clustering(str(tmp_path), "m", 2)[0].tolist() == [3, 0]
    with open(pp, "w") as f:
        f.write("%%MatrixMarket matrix array integer general\n% c\n4 1\n1\n0\n3\n2\n")
    assert read_clustering(str(tmp_path), "m", 4)[0].tolist() == [1, 0, 3, 2]
