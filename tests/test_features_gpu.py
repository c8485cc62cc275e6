"""Matrix features on the device (spops_matrix_features) against the CPU oracle
(orc_features, itself pinned to the reference's MTX extractor): integer-valued features
exactly, the others to 1e-12 relative (different summation order on the device)."""
import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd import features as F
from sparseops_amd import gen

pytestmark = pytest.mark.gpu

INTEGER = [0, 1, 2, 4, 6, 7, 8, 10, 11, 18, 19, 30, 31, 32, 34, 35, 42, 43, 50, 51]

CASES = {
    "pl3000": lambda: gen.powerlaw(3000, lmax=300, seed=2),
    "lap40": lambda: gen.laplacian_2d(40),
    "rmat14": lambda: gen.rmat(14, 16, seed=14),
    "rect700x1900": lambda: gen.random_rect(700, 1900, 20, seed=1),
    "rect100x60": lambda: gen.random_rect(100, 60, 12, seed=4),
    "rect3000x257": lambda: gen.random_rect(3000, 257, 30, seed=5),
    "pl100k": lambda: gen.powerlaw(100000, lmax=4000, seed=7),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_features_match_oracle(name, dev):
    A = CASES[name]()
    want = orc.features(A)
    got = np.array(list(F.features(A, dev).values()))
    np.testing.assert_array_equal(got[INTEGER], want[INTEGER])
    rest = [i for i in range(61) if i not in INTEGER]
    np.testing.assert_allclose(got[rest], want[rest], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("idx,vt", [(torch.int32, torch.float32), (torch.int64, torch.float32),
                                    (torch.int32, torch.float64)])
def test_features_index_and_value_types(idx, vt, dev):
    A = gen.random_rect(2000, 1500, 25, seed=9)
    A32 = gen.Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64))
    want = orc.features(A32 if vt == torch.float32 else A)
    got = F.features_array(torch.tensor(A.rowptr, device=dev).to(idx), torch.tensor(A.col, device=dev).to(idx),
                           torch.tensor(A.val, device=dev).to(vt), A.n_rows, A.n_cols)
    np.testing.assert_array_equal(got[INTEGER], want[INTEGER])
    rest = [i for i in range(61) if i not in INTEGER]
    np.testing.assert_allclose(got[rest], want[rest], rtol=1e-12, atol=1e-300)


def test_features_reject(dev):
    z = torch.zeros(4, dtype=torch.int64, device=dev)
    with pytest.raises(ValueError):
        F.features_array(z, z.int(), torch.zeros(3, device=dev, dtype=torch.float64), 3, 3)
    with pytest.raises(F.L.SpopsError):
        F.features_array(z, z, torch.zeros(3, device=dev, dtype=torch.float64), 0, 3)
