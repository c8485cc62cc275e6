"""Oracle pins of the row-wise SpGEMM, the A * A^T top-k Jaccard front end and the map-based
hierarchical clustering v0, and the library's host clustering v0, on the CPU:

  * orc_topk_jaccard equals the reference's HashSpGEMMTopK output (tests/golden/topk_*, written
    by oracle/_ref/ref_harness) bit for bit, and a brute force over Python sets with
    multiplicities on small matrices (duplicates, empty rows, a rectangular pattern);
  * orc_hc version 2 and spops_hierarchical_clustering_v0 (libsparseops_hip.so, host code)
    equal the reference's hierarchical_clustering_v0 + reordered_dict_to_permutation_and_offsets
    on the pairs of test/test_spgemm_hc.cpp's file round trip (perm, offsets and map roots);
  * orc_spgemm_rows equals the reference's row-wise LeSpGEMM for kernel flags 1, 2 and 3
    (structure exact, row sums / values 1e-12 relative) and scipy's product.
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_bind as orc
from sparseops_amd import lespgemm as LG
from sparseops_amd.gen import Csr, random_rect, rmat

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


TOPK = {"topk_rmat10": (lambda: rmat(10, 16, seed=10), 7, False),
        "topk_rmat10_f32": (lambda: rmat(10, 16, seed=10), 7, True),
        "topk_rmat12_k16": (lambda: rmat(12, 16, seed=12), 16, False),
        "topk_rect900x400": (lambda: random_rect(900, 400, 10, seed=7), 5, False)}


@pytest.mark.parametrize("name", sorted(TOPK))
def test_topk_oracle_matches_reference(name):
    make, k, f32 = TOPK[name]
    g = gold(name)
    A = make()
    T = orc.topk_jaccard(A, k, score_f32=f32)
    np.testing.assert_array_equal(T.rowptr, g["topk_rowptr"])
    np.testing.assert_array_equal(T.col, g["topk_col"])
    np.testing.assert_array_equal(T.val, g["topk_val"])


def _brute_topk(A, k, f32):
    rows = [list(A.col[A.rowptr[i]:A.rowptr[i + 1]]) for i in range(A.n_rows)]
    out = []
    for i, ri in enumerate(rows):
        cand = []
        for j, rj in enumerate(rows):
            if j == i:
                continue
            inter = sum(rj.count(c) for c in ri)        # the product of the 0/1 patterns, multiplicities kept
            if inter == 0:
                continue
            u = len(ri) + len(rj) - inter
            s = (np.float32(inter) / np.float32(u)) if f32 else (inter / u if u else 0.0)
            cand.append((-float(s), j, float(s)))
        cand.sort()
        out.append([(j, s) for _, j, s in cand[:k]])
    return out


@pytest.mark.parametrize("f32", [False, True])
def test_topk_oracle_brute_force(f32):
    rng = np.random.default_rng(11)
    for n_rows, n_cols, per, k in [(40, 25, 5, 3), (60, 60, 7, 7), (30, 80, 3, 50)]:
        rp, ci = [0], []
        for i in range(n_rows):
            m = 0 if i % 9 == 4 else int(rng.integers(1, per + 1))       # some empty rows
            cols = rng.integers(0, n_cols, m)
            if i % 7 == 0 and m:
                cols = np.append(cols, cols[0])                        # a duplicate entry
            ci.extend(cols.tolist())
            rp.append(len(ci))
        A = Csr(n_rows, n_cols, np.array(rp), np.array(ci), np.ones(len(ci)))
        T = orc.topk_jaccard(A, k, score_f32=f32)
        want = _brute_topk(A, k, f32)
        for i in range(n_rows):
            got = list(zip(T.col[T.rowptr[i]:T.rowptr[i + 1]].tolist(), T.val[T.rowptr[i]:T.rowptr[i + 1]].tolist()))
            assert got == want[i], (n_rows, i)


def test_topk_score_is_the_jaccard_index():
    """On duplicate-free rows the score is |A_i & A_j| / |A_i | A_j| exactly (rounded once)."""
    A = rmat(8, 16, seed=3)
    T = orc.topk_jaccard(A, 5)
    sets = [set(A.col[A.rowptr[i]:A.rowptr[i + 1]].tolist()) for i in range(A.n_rows)]
    for i in range(A.n_rows):
        for e in range(T.rowptr[i], T.rowptr[i + 1]):
            j = int(T.col[e])
            exact = Fraction(len(sets[i] & sets[j]), len(sets[i] | sets[j]))
            assert T.val[e] == float(exact)


def digest_of(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _file_pairs(g):
    return g["pair_keys"][:, 0], g["pair_keys"][:, 1], g["pair_scores"]


@pytest.mark.parametrize("name", sorted(TOPK))
def test_hc_v0_oracle_matches_reference(name):
    make, _, f32 = TOPK[name]
    g = gold(name)
    A = make()
    pi, pj, ps = _file_pairs(g)
    perm, offsets = orc.hc(A, pi, pj, ps, cluster_size=int(g["cluster_size"]), version=2, score_f32=f32)
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(offsets, g["offsets"])


@pytest.mark.parametrize("name", sorted(TOPK))
def test_hc_v0_library_matches_reference(name):
    """spops_hierarchical_clustering_v0 is host code of libsparseops_hip.so (no GPU needed)."""
    make, _, f32 = TOPK[name]
    g = gold(name)
    A = make()
    pi, pj, ps = _file_pairs(g)
    perm, offsets, roots = LG.hierarchical_clustering_v0(A, pi, pj, ps.astype(np.float32) if f32 else ps,
                                                         int(g["cluster_size"]))
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(offsets, g["offsets"])
    np.testing.assert_array_equal(roots, g["roots"])


def test_hc_v0_reversed_keys_and_self_pair():
    """Keys held as (i, j) with i > j are queued as (min, max) and still block the rescoring of
    that pair; a self pair (i, i) of a root doubles its size (the reference's behaviour): the
    library and the oracle agree on such inputs."""
    A = rmat(8, 16, seed=8)
    pi = np.array([5, 3, 7, 2, 9, 9], np.int64)
    pj = np.array([3, 5, 2, 7, 9, 1], np.int64)
    ps = np.array([0.9, 0.8, 0.7, 0.6, 0.95, 0.5])
    for cs in (2, 3, 8):
        perm_o, off_o = orc.hc(A, pi, pj, ps, cluster_size=cs, version=2)
        perm_l, off_l, _ = LG.hierarchical_clustering_v0(A, pi, pj, ps, cs)
        np.testing.assert_array_equal(perm_l, perm_o)
        np.testing.assert_array_equal(off_l, off_o)


ROWWISE = {"rowwise_rmat10": (lambda: rmat(10, 16, seed=10), lambda: rmat(10, 16, seed=10)),
           "rowwise_rect_f64": (lambda: random_rect(300, 200, 12, seed=5), lambda: random_rect(200, 500, 9, seed=6))}


@pytest.mark.parametrize("name", sorted(ROWWISE))
def test_rowwise_oracle_matches_reference(name):
    """Kernel flags 1 and 2 on structure and values, flag 3 on structure: the reference's SPA
    kernel never clears its dense accumulator before first use (the clearing loop of
    src/spgemm_array.cpp:453-457 is commented out), so its values start from whatever the
    allocation held (15 of 8096 entries differ in the rectangular case; the R-MAT case happened
    to get zeroed memory)."""
    g = gold(name)
    A, B = ROWWISE[name][0](), ROWWISE[name][1]()
    C = orc.spgemm_rows(A, B)
    rows = np.repeat(np.arange(A.n_rows), np.diff(C.rowptr))
    rs = np.bincount(rows, weights=C.val, minlength=A.n_rows)
    np.testing.assert_array_equal(C.rowptr, g["k3_rowptr"])
    assert digest_of(C.col) == str(g["k3_col_digest"])
    for kern in (1, 2):
        np.testing.assert_array_equal(C.rowptr, g[f"k{kern}_rowptr"])
        assert digest_of(C.col) == str(g[f"k{kern}_col_digest"])
        np.testing.assert_allclose(rs, g[f"k{kern}_rowsum"], rtol=1e-12, atol=1e-12 * np.abs(rs).max())
        if f"k{kern}_col" in g:
            np.testing.assert_array_equal(C.col, g[f"k{kern}_col"])
            np.testing.assert_allclose(C.val, g[f"k{kern}_val"], rtol=1e-12, atol=0)


def test_rowwise_oracle_matches_scipy():
    """Structure: the product of the 0/1 patterns (zero sums and explicit zeros kept, which
    scipy would drop); values: scipy's product, densely."""
    for A, B in [(rmat(9, 8, seed=1), rmat(9, 8, seed=2)), (random_rect(50, 400, 6, seed=1), random_rect(400, 30, 4, seed=2))]:
        C = orc.spgemm_rows(A, B)
        sa = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
        sb = sp.csr_matrix((B.val, B.col, B.rowptr), shape=(B.n_rows, B.n_cols))
        pa = sp.csr_matrix((np.ones(A.nnz), A.col, A.rowptr), shape=sa.shape)
        pb = sp.csr_matrix((np.ones(B.nnz), B.col, B.rowptr), shape=sb.shape)
        P = (pa @ pb).tocsr()
        P.sort_indices()
        np.testing.assert_array_equal(C.rowptr, P.indptr)
        np.testing.assert_array_equal(C.col, P.indices)
        got = sp.csr_matrix((C.val, C.col, C.rowptr), shape=(A.n_rows, B.n_cols)).toarray()
        want = (sa @ sb).toarray()
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
