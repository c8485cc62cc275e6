"""The benchmark configurations at their full sizes, through the C ABI, against the CPU oracle
and the reference's own fixtures:

  * power-law n = 2,097,152 (mean row length 15.5, max 8192): SELL-C-sigma (C = 32,
    sigma = 1024) and CSR SpMV in FP64, FP32 and BF16 value storage; the SELL reorder and
    chunk_len against the oracle's layout;
  * LeSpGEMM C = A * A on R-MAT scales 14 and 16 (every output compared by digest and row
    sums with the reference's) and on R-MAT scale 20 (clustering digests and two sampled
    cluster panels' checksums against the reference's LeSpGEMM_VLength);
  * the batched SpMM of 512 band matrices (every matrix, every output);
  * the differentiable SpMM on R-MAT scale 20, edge factor 32 (Y, dX and dvals on the whole
    matrix).

Tolerances are those of the per-kernel tests: SpMV conditioned error 1e-13 (FP64) and 4e-7
(FP32 / BF16 storage, against the exact product of the stored values); LeSpGEMM structure
exact, row sums 1e-12, panel value sums 1e-10 of the sum of |values| (sums of 5e7-2e8 terms in
different orders); SpMM / SDDMM 2e-6 of sum |a x| (FP32 accumulation)."""
import hashlib
import os

import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd import autograd as AG
from sparseops_amd import lespgemm as LG
from sparseops_amd import spmv as S
from sparseops_amd.batch import BatchSpMM
from sparseops_amd.gen import Csr, band_batch, powerlaw, rmat, rmat_symnorm
from test_batch_gpu import _check as batch_check
from test_batch_gpu import _fill as batch_fill
from test_spmv_gpu import _check as spmv_check
from test_spmv_gpu import _y0

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def log(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ power-law SpMV, n = 2^21
@pytest.fixture(scope="module")
def pl2m():
    A = powerlaw(1 << 21, seed=0)
    log(f"power-law n={A.n_rows} nnz={A.nnz}")
    return A


@pytest.mark.timeout(900)
@pytest.mark.parametrize("vt", ["f32", "bf16", "f64"])
@pytest.mark.parametrize("fmt", ["sell", "csr"])
def test_powerlaw_2m_spmv(pl2m, fmt, vt, dev):
    A = pl2m
    log(f"power-law SpMV {fmt} {vt}")
    if fmt == "sell":
        M = S.csr_to_sell_c_sigma(A, 1024, 32, vt)
        reorder, cl, _, _, _ = orc.sell_layout(A, 1024, 32, 4 if vt == "f64" else 8)
        np.testing.assert_array_equal(M.layout(0), reorder)
        np.testing.assert_array_equal(M.layout(1), cl)
    else:
        M = S.csr(A, vt)
    assert M.sweep is not None          # both run on the swept layout at this size
    x = np.random.default_rng(0).uniform(-1.0, 1.0, A.n_cols)
    spmv_check(A, M, x, np.zeros(A.n_rows), vt, alpha=1.0, beta=0.0, dev=dev)
    spmv_check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    M.close()


# --------------------------------------------------------------------------- LeSpGEMM, R-MAT
@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,scale", [("lsh_rmat14_fp", 14), ("lsh_rmat16_fp", 16)])
def test_lespgemm_rmat_matches_reference_fingerprints(name, scale, dev):
    g = gold(name)
    A = rmat(scale, 16, seed=scale)
    assert digest(A.rowptr, A.col, A.val) == str(g["input_digest"])
    log(f"LeSpGEMM R-MAT {scale}: nnz(A)={A.nnz}")
    out = LG.lespgemm(A, A, k=64, bands=16, seed=7, cluster_size=8, device=dev)
    assert digest(out["perm"]) == str(g["perm_digest"])
    assert digest(out["offsets"]) == str(g["offsets_digest"])
    assert out["flop"] == int(g["flops"][0])
    assert int(out["C_rowptr"][-1]) == int(g["nnz"])
    assert digest(out["C_rowptr"]) == str(g["C_rowptr_digest"])
    assert digest(np.asarray(out["C_col"], np.int64)) == str(g["C_col_digest"])
    rows = np.repeat(np.arange(A.n_rows), np.diff(out["C_rowptr"]))
    rs = np.bincount(rows, weights=out["C_val"], minlength=A.n_rows)
    b = int(g["block"])
    np.testing.assert_allclose(rs.reshape(-1, b).sum(axis=1), g["C_rowsum"], rtol=1e-12, atol=0)


@pytest.mark.timeout(900)
def test_lespgemm_rmat20_panels_match_reference(dev):
    """R-MAT scale 20 (config size): the GPU clustering equals the reference's (digests), and
    two cluster panels of about 2 % of the flop each equal the reference's LeSpGEMM_VLength
    output on the same clusters (C's column count, value count, column-id sum and nonzero count
    exact; value sum within 1e-10 of the sum of |values|)."""
    g = gold("lsh_rmat20_panels")
    A = rmat(20, 16, seed=0)
    assert digest(A.rowptr, A.col, A.val) == str(g["input_digest"])
    log(f"LeSpGEMM R-MAT 20: nnz(A)={A.nnz}")
    dA = LG.DeviceCsr(A, dev)
    pi, pj, ps = LG.lsh_candidate_pairs(dA, 64, 16, 7, constrain=True)
    perm, offsets = LG.hierarchical_clustering(A, pi, pj, ps, 8, version=0)
    assert digest(perm) == str(g["perm_digest"]) and digest(offsets) == str(g["offsets_digest"])
    Acl = LG.ClusterMatrix.build(dA, perm, offsets)
    plan = LG.SpgemmPlan(Acl, dA)
    picks = LG.sample_ranges(LG.cluster_flops(A, A.rowptr, perm, offsets), float(g["share"]))
    np.testing.assert_array_equal(np.array(picks), g["ranges"])
    for t, (c0, c1) in enumerate(picks):
        log(f"  panel {c0}:{c1}")
        P = plan.numeric(c0, c1)
        cs = P.checksum()
        assert (P.nnzc, P.nnzv) == (int(g["nnzc"][t]), int(g["nnzv"][t]))
        assert cs["colid_sum"] == int(g["colid_sum"][t]) and cs["nonzeros"] == int(g["nonzeros"][t])
        assert abs(cs["sum"] - float(g["sum"][t])) <= 1e-10 * float(g["sum_abs"][t])
        P.close()
    plan.close()
    Acl.close()


# ------------------------------------------------------------------ batched SpMM, 512 matrices
@pytest.mark.timeout(900)
def test_batch_512_band_matrices(dev):
    mats, ks = band_batch(512, seed=0)
    log(f"batch: {len(mats)} matrices, nnz {sum(A.nnz for A in mats)}")
    B = BatchSpMM(mats, ks, tune=True)
    X, Y = B.new_x(), B.new_y()
    batch_fill(B, X, 7)
    Y.fill_(float("nan"))
    B.spmm(X, Y)
    torch.cuda.synchronize()
    batch_check(B, mats, X, Y)
    B.close()


# ------------------------------------------------ differentiable SpMM, R-MAT 20, edge factor 32
def _close(got, want, scale, tol, what):
    err = np.abs(got - want) / (scale + 1e-300)
    assert np.all(err <= tol), f"{what}: max scaled error {err.max():.3e} > {tol}"


@pytest.mark.timeout(900)
def test_autograd_rmat20_ef32(dev):
    """Y = A X, dX = A^T G and dvals = SDDMM(G, X) of the layer (FP32 symmetric-normalised
    values, BF16 X and G, FP32 accumulation) on the whole matrix against the CPU oracle on
    the same rounded inputs."""
    A = rmat_symnorm(20, 32, seed=0)
    log(f"autograd R-MAT 20 ef 32: nnz={A.nnz}")
    pat = AG.SparsePattern(A, "f32")
    g = torch.Generator(device="cpu").manual_seed(5)
    vals = torch.tensor(A.val, device=dev, requires_grad=True)
    X = torch.randn(A.n_cols, 64, generator=g).to(torch.bfloat16).to(dev).requires_grad_()
    G = torch.randn(A.n_rows, 64, generator=g).to(torch.bfloat16).to(dev)
    Y = AG.spmm(pat, vals, X, out_dtype=torch.float32)
    Y.backward(G.float())
    torch.cuda.synchronize()
    Xn, Gn = X.detach().double().cpu().numpy(), G.double().cpu().numpy()
    A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
    absA = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A64.val))
    _close(Y.detach().double().cpu().numpy(), orc.spmm(A64, Xn), orc.spmm(absA, np.abs(Xn)), 2e-6, "Y")
    del Y
    trp, tc, tv, _ = orc.transpose(A64)
    want = orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, tv), Gn)
    scale = orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, np.abs(tv)), np.abs(Gn))
    _close(X.grad.double().cpu().numpy(), want, scale, 2 ** -8 + 2e-6, "dX")     # dX rounded to BF16
    del want, scale
    _close(vals.grad.double().cpu().numpy(), orc.sddmm(A64, Gn, Xn), orc.sddmm(A64, np.abs(Gn), np.abs(Xn)), 2e-6,
           "dvals")
    pat.close()
