"""The approximate GPU clustering (spops_cluster_matching, no reference counterpart) against its
oracle (orc_cluster_matching, pinned in tests/test_cluster_matching_cpu.py): permutation, offsets
and round count bit-identical on the LSH pairs of R-MAT matrices (FP64 and FP32 scores) and on
random graphs with tied scores and invalid pairs; the LeSpGEMM pipeline with it (hc_v = 3) gives
the product scipy gives."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle_bind as orc
from sparseops_amd import lespgemm as LG
from sparseops_amd.gen import rmat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,cs,f32", [(10, 8, False), (12, 8, False), (12, 4, True), (14, 16, False)])
def test_matching_matches_oracle_on_lsh_pairs(scale, cs, f32, dev):
    A = rmat(scale, 16, seed=scale)
    dA = LG.DeviceCsr(A, dev)
    print(f"matching R-MAT {scale} cs={cs} f32={f32}", flush=True)
    pi, pj, ps = LG.lsh_candidate_pairs(dA, 64, 16, 7, constrain=True, score_f32=f32)
    perm, off, rounds, npairs = LG.cluster_rows_gpu(dA, 64, 16, 7, True, cs, score_f32=f32)
    assert npairs == len(pi)
    want = orc.cluster_matching(A.n_rows, pi, pj, ps.astype(np.float64), cs)
    np.testing.assert_array_equal(perm, want[0])
    np.testing.assert_array_equal(off, want[1])
    assert rounds == want[2] and rounds < 256


@pytest.mark.parametrize("seed", range(4))
def test_matching_matches_oracle_on_random_graphs(seed, dev):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 3000))
    m = int(rng.integers(0, 6 * n))
    pi = rng.integers(-2, n + 2, m)
    pj = rng.integers(0, n, m)
    sc = rng.choice([0.0, 0.1, 0.5, 0.5, 0.9], m)
    for cs in (2, 8):
        got = LG.cluster_matching(n, torch.tensor(pi, device=dev), torch.tensor(pj, device=dev),
                                  torch.tensor(sc, device=dev), cs)
        want = orc.cluster_matching(n, pi, pj, sc, cs)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
        assert got[2] == want[2]


def test_pipeline_with_gpu_clustering(dev):
    A = rmat(12, 16, seed=12)
    out = LG.lespgemm(A, A, hc_v=3, device=dev)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
    W = (S @ S).tocsr()
    W.sort_indices()
    np.testing.assert_array_equal(out["C_rowptr"], W.indptr)
    np.testing.assert_array_equal(out["C_col"], W.indices)
    np.testing.assert_allclose(out["C_val"], W.data, rtol=1e-12, atol=0)
    assert out["times"]["hc_rounds"] > 0
