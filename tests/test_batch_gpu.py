"""Batched SpMM (one persistent launch over many FP32 matrices, each with its own k) through
the C ABI, against the CPU oracle's SpMM (orc_spmm_csr) matrix by matrix.

Tolerance: FP32 values and X, FP32 accumulation: 2e-6 relative to sum |a x| per output,
against the exact (FP64) product of the values as stored."""
import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd.batch import BatchSpMM
from sparseops_amd.gen import Csr, band_batch, band_matrix, powerlaw

pytestmark = pytest.mark.gpu


def _mats():
    """Band matrices of both formats' k ranges (1..70), a power-law matrix with a long row,
    an empty matrix, and one whose rows are all empty."""
    rng = np.random.default_rng(1)
    mats, ks = band_batch(count=20, nmin=100, nmax=4000, wmax=40, kmax=40, seed=3)
    ks[0], ks[1], ks[2] = 1, 2, 70
    mats.append(powerlaw(3000, lmax=1500, seed=5))
    ks.append(3)
    mats.append(Csr(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)))
    ks.append(4)
    mats.append(Csr(50, 30, np.zeros(51, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)))
    ks.append(5)
    mats.append(band_matrix(777, 3, 2, rng))
    ks.append(33)
    return [Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.asarray(A.val, np.float32)) for A in mats], ks


@pytest.fixture(scope="module")
def batch():
    return _mats()


def _fill(B, X, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    X.copy_(torch.randn(X.numel(), generator=g).to(X.device))


def _check(B, mats, X, Y, alpha=1.0, beta=0.0, Y0=None):
    Xh, Yh = X.double().cpu(), Y.double().cpu()
    for i, A in enumerate(mats):
        if A.n_rows == 0:
            continue
        k = B.ks[i]
        Xi = B.x_view(Xh, i).numpy()
        got = B.y_view(Yh, i).numpy()
        A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
        want = alpha * orc.spmm(A64, Xi)
        scale = abs(alpha) * orc.spmm(Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A64.val)), np.abs(Xi))
        if Y0 is not None:
            y0 = B.y_view(Y0.double().cpu(), i).numpy()
            want = want + beta * y0
            scale = scale + abs(beta) * np.abs(y0)
        err = np.abs(got - want) / (scale + 1e-30)
        assert np.all(err <= 2e-6), f"matrix {i} ({B.formats()[i]}, k={k}): max scaled error {err.max():.3e}"


@pytest.mark.parametrize("tune", [False, True])
def test_batch_matches_oracle(batch, tune, dev):
    mats, ks = batch
    B = BatchSpMM(mats, ks, tune=tune)
    assert B.flop == sum(2 * A.nnz * k for A, k in zip(mats, ks))
    # packed blocks: 16-byte aligned, row stride ld >= k, disjoint, inside x_len / y_len
    for off, dim, total in ((B.x_off, 1, B.x_len), (B.y_off, 0, B.y_len)):
        spans = sorted((int(off[i]), int(off[i]) + B.shapes[i][dim] * int(B.ld[i])) for i in range(B.n))
        assert all(a % 4 == 0 for a, _ in spans) and all(B.ld[i] >= ks[i] for i in range(B.n))
        assert all(spans[j][1] <= spans[j + 1][0] for j in range(len(spans) - 1)) and spans[-1][1] <= total
    if not tune:      # the default rule: row chunks for k <= 64, CSR lane groups beyond
        assert [f == "chunks" for f in B.formats()] == [k <= 64 for k in ks]
    X, Y = B.new_x(), B.new_y()
    _fill(B, X, 0)
    Y.fill_(float("nan"))
    B.spmm(X, Y)
    torch.cuda.synchronize()
    _check(B, mats, X, Y)
    B.close()


def test_batch_alpha_beta(batch, dev):
    mats, ks = batch
    B = BatchSpMM(mats, ks, tune=False)
    X, Y = B.new_x(), B.new_y()
    _fill(B, X, 1)
    _fill(B, Y, 2)
    Y0 = Y.clone()
    B.spmm(X, Y, alpha=-0.5, beta=1.25)
    torch.cuda.synchronize()
    _check(B, mats, X, Y, -0.5, 1.25, Y0)


def test_batch_graph_replays(batch, dev):
    """Capture once, replay three times with eager work (new X, a torch reduction, a host
    synchronisation) between replays; every replay matches the eager launch on its X."""
    mats, ks = batch
    B = BatchSpMM(mats, ks, tune=True)
    X, Y, Ye = B.new_x(), B.new_y(), B.new_y()
    s = torch.cuda.Stream()
    _fill(B, X, 3)
    torch.cuda.synchronize()
    B.capture(X, Y, s.cuda_stream)
    for rep in range(3):
        _fill(B, X, 10 + rep)
        Y.fill_(float("nan"))
        _ = X.abs().sum().item()
        torch.cuda.synchronize()
        B.replay(s.cuda_stream)
        s.synchronize()
        B.spmm(X, Ye)
        torch.cuda.synchronize()
        assert all(torch.equal(B.y_view(Y, i), B.y_view(Ye, i)) for i in range(B.n)), \
            f"replay {rep} differs from the eager launch"
        _check(B, mats, X, Y)


def test_batch_rejects(dev):
    A = band_matrix(100, 2, 0, np.random.default_rng(0))
    with pytest.raises(ValueError):
        BatchSpMM([A], [1, 2])
    bad = Csr(A.n_rows, A.n_cols, A.rowptr, A.col.copy(), A.val)
    bad.col[5] = A.n_cols
    with pytest.raises(Exception):
        BatchSpMM([bad], [2])
    with pytest.raises(Exception):
        BatchSpMM([A], [0])
    B = BatchSpMM([A], [2])
    with pytest.raises(ValueError):
        B.spmm(torch.zeros(3, device=dev), B.new_y())
    with pytest.raises(ValueError):                          # not 16-byte aligned
        B.spmm(torch.zeros(B.x_len + 1, device=dev)[1:], B.new_y())
    with pytest.raises(Exception):
        B.replay(torch.cuda.current_stream().cuda_stream)     # nothing captured yet
