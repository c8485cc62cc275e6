"""SpMM, SDDMM, transpose and the differentiable SpMM (through the C ABI) against the CPU
oracle (orc_spmm_csr, orc_sddmm_csr, orc_csr_transpose) and FP64 finite differences.

Tolerances: FP64 data, FP64 accumulation: 1e-13 relative to sum |a x| per output; FP32 /
BF16 data with FP32 accumulation: 2e-6 of that scale against the exact product of the
values as stored (BF16 inputs are rounded on the host first, so only the accumulation is
measured)."""
import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd import autograd as AG
from sparseops_amd import spmv as S
from sparseops_amd.gen import Csr, powerlaw, rmat, rmat_symnorm

pytestmark = pytest.mark.gpu


def _edge(seed=4):
    """Rows of every SpMM path: empty, short, long (> 2048 entries)."""
    rng = np.random.default_rng(seed)
    n = 5000
    lens = rng.integers(0, 30, n)
    lens[::50] = 0
    lens[7], lens[1500] = 2600, 4100
    cols = [np.sort(rng.choice(n, size=L, replace=False)) for L in lens]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.concatenate(cols)
    return Csr(n, n, rp.astype(np.int32), col.astype(np.int32), rng.standard_normal(col.size))


def _bf16(a):
    return torch.tensor(a, dtype=torch.float64).to(torch.bfloat16).to(torch.float64).numpy()


def _close(got, want, scale, tol):
    err = np.abs(got - want) / (scale + 1e-300)
    assert np.all(err <= tol), f"max scaled error {err.max():.3e} > {tol}"


@pytest.fixture(scope="module")
def mats():
    return {"edge": _edge(), "pl": powerlaw(5000, lmax=800, seed=3), "rmat": rmat(11, 8, seed=2, index_dtype=np.int32)}


@pytest.mark.parametrize("k", [1, 3, 8, 17, 64, 100])
@pytest.mark.parametrize("name", ["edge", "pl", "rmat"])
def test_spmm_f64(mats, name, k, dev):
    A = mats[name]
    rng = np.random.default_rng(k)
    X = rng.standard_normal((A.n_cols, k))
    Y0 = rng.standard_normal((A.n_rows, k))
    M = S.csr(A, "f64")
    Xt = torch.tensor(X, device=dev)
    Yt = torch.tensor(Y0, device=dev)
    M.spmm(Xt, Yt, 0.75, -0.5)
    want = 0.75 * orc.spmm(A, X) - 0.5 * Y0
    Aabs = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A.val))
    scale = 0.75 * orc.spmm(Aabs, np.abs(X)) + 0.5 * np.abs(Y0)
    _close(Yt.cpu().numpy(), want, scale, 1e-13)


@pytest.mark.parametrize("xt,yt", [(torch.float32, torch.float32), (torch.bfloat16, torch.float32),
                                   (torch.bfloat16, torch.bfloat16), (torch.float32, torch.float64)])
def test_spmm_mixed_types(mats, xt, yt, dev):
    A = mats["edge"]
    k = 64
    rng = np.random.default_rng(5)
    X = rng.standard_normal((A.n_cols, k))
    M = S.csr(A, "f32")
    Xt = torch.tensor(X, device=dev).to(xt)
    Yt = torch.zeros(A.n_rows, k, dtype=yt, device=dev)
    M.spmm(Xt, Yt)
    Xr = Xt.double().cpu().numpy()
    Ar = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64))
    want = orc.spmm(Ar, Xr)
    scale = orc.spmm(Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(Ar.val)), np.abs(Xr))
    tol = 2e-6 if yt != torch.bfloat16 else 2 ** -8 + 2e-6      # BF16 output: its own rounding
    _close(Yt.double().cpu().numpy(), want, scale, tol)


@pytest.mark.parametrize("k", [1, 16, 64, 70, 512, 520])
@pytest.mark.parametrize("gt,xt", [(torch.float64, torch.float64), (torch.float32, torch.float32),
                                   (torch.bfloat16, torch.bfloat16), (torch.float32, torch.bfloat16)])
@pytest.mark.parametrize("name", ["rmat", "edge"])
def test_sddmm(mats, name, k, gt, xt, dev):
    """By rows (each G row loaded once per chunk of at most 64 entries of a row)
    when k is a multiple of X's 16-byte vector and at most 64 vectors; entry-parallel otherwise
    (k = 1, and k = 70 / 512 / 520 where the vector does not divide k or 64 vectors fall short)."""
    A = mats[name]
    rng = np.random.default_rng(k + 1)
    G = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).to(gt)
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(xt)
    M = S.csr(A, "f64")
    out = torch.zeros(A.nnz, dtype=torch.float64 if gt == torch.float64 else torch.float32, device=dev)
    M.sddmm(G, X, out)
    Gn, Xn = G.double().cpu().numpy(), X.double().cpu().numpy()
    want = orc.sddmm(A, Gn, Xn)
    scale = orc.sddmm(A, np.abs(Gn), np.abs(Xn))
    _close(out.double().cpu().numpy(), want, scale, 1e-13 if gt == torch.float64 else 2e-6)


@pytest.mark.parametrize("vt,xt", [("f64", torch.float64), ("f32", torch.float32), ("f32", torch.bfloat16),
                                   ("bf16", torch.bfloat16)])
@pytest.mark.parametrize("name", ["edge", "rmat"])
def test_spmm_values_in_place(mats, name, vt, xt, dev):
    """spops_spmm_values reads the values from the caller's array: the same product as
    set_values + spmm on those values, bit for bit, and the stored values stay as they were."""
    A = mats[name]
    k = 64
    rng = np.random.default_rng(7)
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[vt]
    vals = torch.tensor(rng.standard_normal(A.nnz), device=dev).to(tdt)
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(xt)
    M = S.csr(A, vt)
    yt = torch.float64 if vt == "f64" else torch.float32
    Y0 = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).to(yt)
    Y1 = Y0.clone()
    M.spmm_values(vals, X, Y1, 0.5, 0.25)
    Ys = Y0.clone()
    M.spmm(X, Ys, 0.5, 0.25)                      # the stored values: A.val
    R = S.csr(A, vt)
    R.set_values(vals)
    Y2 = Y0.clone()
    R.spmm(X, Y2, 0.5, 0.25)
    torch.cuda.synchronize()
    assert torch.equal(Y1, Y2)
    Ystored = Y0.clone()
    S.csr(A, vt).spmm(X, Ystored, 0.5, 0.25)
    assert torch.equal(Ys, Ystored)
    # a transposed handle reads the values (in A's entry order) through its permutation
    AT, AT2 = M.transpose(), S.csr(A, vt).transpose()
    AT2.set_values(vals)
    Gm = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).to(xt)
    D0 = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(yt)
    D1, D2 = D0.clone(), D0.clone()
    AT.spmm_values(vals, Gm, D1, 1.5, -0.5)
    AT2.spmm(Gm, D2, 1.5, -0.5)
    torch.cuda.synchronize()
    assert torch.equal(D1, D2)
    with pytest.raises(ValueError):
        M.spmm_values(vals[:-1], X, Y1)
    other = torch.float32 if vt != "f32" else torch.float64
    with pytest.raises(ValueError):
        M.spmm_values(vals.to(other), X, Y1)


@pytest.mark.parametrize("name", ["edge", "rmat"])
def test_transpose_and_value_updates(mats, name, dev):
    A = mats[name]
    M = S.csr(A, "f64")
    T = M.transpose()
    trp, tc, tv, tp = orc.transpose(A)
    assert (T.n_rows, T.n_cols, T.nnz) == (A.n_cols, A.n_rows, A.nnz)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((A.n_rows, 5))
    Y = torch.zeros(A.n_cols, 5, dtype=torch.float64, device=dev)
    T.spmm(torch.tensor(X, device=dev), Y)
    At = Csr(A.n_cols, A.n_rows, trp, tc, tv)
    np.testing.assert_allclose(Y.cpu().numpy(), orc.spmm(At, X), rtol=1e-12, atol=1e-12)
    # new values (in A's entry order) reach both handles
    v2 = rng.standard_normal(A.nnz)
    vt = torch.tensor(v2, device=dev)
    M.set_values(vt)
    T.set_values(vt)
    Y.zero_()
    T.spmm(torch.tensor(X, device=dev), Y)
    np.testing.assert_allclose(Y.cpu().numpy(), orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, v2[tp]), X), rtol=1e-12,
                               atol=1e-12)
    X2 = rng.standard_normal((A.n_cols, 5))
    Y2 = torch.zeros(A.n_rows, 5, dtype=torch.float64, device=dev)
    M.spmm(torch.tensor(X2, device=dev), Y2)
    np.testing.assert_allclose(Y2.cpu().numpy(), orc.spmm(Csr(A.n_rows, A.n_cols, A.rowptr, A.col, v2), X2),
                               rtol=1e-12, atol=1e-12)


def test_gradcheck_fp64_finite_differences(dev):
    """torch.autograd.gradcheck: analytic dvals and dX against FP64 central differences."""
    A = rmat(6, 4, seed=11, index_dtype=np.int32)
    pat = AG.SparsePattern(A, "f64")
    rng = np.random.default_rng(0)
    vals = torch.tensor(rng.standard_normal(A.nnz), device=dev, requires_grad=True)
    X = torch.tensor(rng.standard_normal((A.n_cols, 3)), device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda v, x: AG.spmm(pat, v, x), (vals, X), eps=1e-6, atol=1e-9, rtol=1e-7)


def test_backward_matches_cpu_reference_bf16(dev):
    """The layer's form: FP32 symmetric-normalised values, BF16 X and G, FP32 accumulation.
    dX = A^T G and dvals = SDDMM(G, X) against the CPU oracle on the same rounded inputs."""
    A = rmat_symnorm(12, 16, seed=1)
    pat = AG.SparsePattern(A, "f32")
    rng = np.random.default_rng(2)
    vals = torch.tensor(A.val, device=dev, requires_grad=True)
    X = torch.tensor(rng.standard_normal((A.n_cols, 64)), device=dev).to(torch.bfloat16).requires_grad_()
    Y = AG.spmm(pat, vals, X, out_dtype=torch.float32)
    G = torch.tensor(rng.standard_normal((A.n_rows, 64)), device=dev).to(torch.bfloat16)
    Y.backward(G.float())
    Xn, Gn = X.detach().double().cpu().numpy(), G.double().cpu().numpy()
    A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
    absA = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A64.val))
    # forward
    _close(Y.detach().double().cpu().numpy(), orc.spmm(A64, Xn), orc.spmm(absA, np.abs(Xn)), 2e-6)
    # dX = A^T G (rounded to BF16 on output: X's dtype)
    trp, tc, tv, tp = orc.transpose(A64)
    At = Csr(A.n_cols, A.n_rows, trp, tc, tv)
    want = orc.spmm(At, Gn)
    scale = orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, np.abs(tv)), np.abs(Gn))
    assert X.grad.dtype == torch.bfloat16
    _close(X.grad.double().cpu().numpy(), want, scale, 2 ** -8 + 2e-6)
    # dvals
    _close(vals.grad.double().cpu().numpy(), orc.sddmm(A64, Gn, Xn), orc.sddmm(A64, np.abs(Gn), np.abs(Xn)), 2e-6)


def test_backward_bf16_output_matches_cpu_reference(dev):
    """The layer with BF16 output: Y rounded once to BF16 from FP32 sums, so its gradient G
    arrives in BF16 and the fused backward gathers BF16 rows of G.  Y, dX = A^T G and
    dvals = SDDMM(G, X) against the CPU oracle on the same rounded inputs."""
    A = rmat_symnorm(12, 16, seed=5)
    pat = AG.SparsePattern(A, "f32")
    rng = np.random.default_rng(6)
    vals = torch.tensor(A.val, device=dev, requires_grad=True)
    X = torch.tensor(rng.standard_normal((A.n_cols, 64)), device=dev).to(torch.bfloat16).requires_grad_()
    Y = AG.spmm(pat, vals, X, out_dtype=torch.bfloat16)
    assert Y.dtype == torch.bfloat16
    G = torch.tensor(rng.standard_normal((A.n_rows, 64)), device=dev).to(torch.bfloat16)
    Y.backward(G)
    Xn, Gn = X.detach().double().cpu().numpy(), G.double().cpu().numpy()
    A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
    absA = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A64.val))
    _close(Y.detach().double().cpu().numpy(), orc.spmm(A64, Xn), orc.spmm(absA, np.abs(Xn)), 2 ** -8 + 2e-6)
    trp, tc, tv, tp = orc.transpose(A64)
    want = orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, tv), Gn)
    scale = orc.spmm(Csr(A.n_cols, A.n_rows, trp, tc, np.abs(tv)), np.abs(Gn))
    assert X.grad.dtype == torch.bfloat16 and vals.grad.dtype == torch.float32
    _close(X.grad.double().cpu().numpy(), want, scale, 2 ** -8 + 2e-6)
    _close(vals.grad.double().cpu().numpy(), orc.sddmm(A64, Gn, Xn), orc.sddmm(A64, np.abs(Gn), np.abs(Xn)), 2e-6)
    # the same G as FP32 values takes the FP32-gather form of the fused pass: equal to within
    # FP32 summation order
    vals2 = torch.tensor(A.val, device=dev, requires_grad=True)
    X2 = X.detach().clone().requires_grad_()
    AG.spmm(pat, vals2, X2, out_dtype=torch.float32).backward(G.float())
    _close(vals2.grad.double().cpu().numpy(), vals.grad.double().cpu().numpy(),
           orc.sddmm(A64, np.abs(Gn), np.abs(Xn)), 2e-6)
    _close(X2.grad.double().cpu().numpy(), X.grad.double().cpu().numpy(), scale, 2 ** -7)


@pytest.mark.parametrize("k", [1, 5, 8, 33])
def test_spmm_writes_nothing_past_y(k, dev):
    """Y sits inside a larger buffer whose tail must stay untouched (rows past n_rows in the
    last workgroup write nothing)."""
    A = rmat(9, 4, seed=6, index_dtype=np.int32)
    M = S.csr(A, "f64")
    buf = torch.full((A.n_rows * k + 4096,), 7.0, dtype=torch.float64, device=dev)
    Y = buf[:A.n_rows * k].view(A.n_rows, k)
    X = torch.ones(A.n_cols, k, dtype=torch.float64, device=dev)
    M.spmm(X, Y)
    torch.cuda.synchronize()
    assert torch.all(buf[A.n_rows * k:] == 7.0)
    want = np.repeat(np.add.reduceat(np.r_[A.val, 0.0], np.r_[A.rowptr[:-1]])[:, None], k, axis=1)
    want[np.diff(A.rowptr) == 0] = 0.0
    np.testing.assert_allclose(Y.cpu().numpy(), want, rtol=1e-12, atol=1e-12)


def test_spmm_rejects_other_formats_and_shapes(mats, dev):
    A = mats["pl"]
    M = S.csr_to_sell_c_sigma(A, 1024, 32, "f32")
    X = torch.zeros(A.n_cols, 4, device=dev)
    Y = torch.zeros(A.n_rows, 4, device=dev)
    with pytest.raises(S.L.SpopsError):
        M.spmm(X, Y)
    C = S.csr(A, "f32")
    with pytest.raises(ValueError):
        C.spmm(X[:-1], Y)


@pytest.mark.parametrize("offset", [0, 1, 2])
def test_spmm_sddmm_unaligned_views(mats, offset, dev):
    """X, Y and G as views at an element offset: the 16-byte vector path is used only when
    the pointers allow it; every offset gives the same products."""
    A = mats["edge"]
    k = 32
    rng = np.random.default_rng(offset)
    Xn = rng.standard_normal((A.n_cols, k)).astype(np.float32)
    Gn = rng.standard_normal((A.n_rows, k)).astype(np.float32)
    xb = torch.zeros(A.n_cols * k + 8, device=dev)
    xb[offset:offset + A.n_cols * k] = torch.tensor(Xn.ravel(), device=dev)
    X = xb[offset:offset + A.n_cols * k].view(A.n_cols, k)
    yb = torch.zeros(A.n_rows * k + 8, device=dev)
    Y = yb[offset:offset + A.n_rows * k].view(A.n_rows, k)
    gb = torch.zeros(A.n_rows * k + 8, device=dev)
    gb[offset:offset + A.n_rows * k] = torch.tensor(Gn.ravel(), device=dev)
    G = gb[offset:offset + A.n_rows * k].view(A.n_rows, k)
    M = S.csr(A, "f32")
    M.spmm(X, Y)
    out = torch.zeros(A.nnz, device=dev)
    M.sddmm(G, X, out)
    Ar = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64))
    absA = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(Ar.val))
    X64, G64 = Xn.astype(np.float64), Gn.astype(np.float64)
    _close(Y.double().cpu().numpy(), orc.spmm(Ar, X64), orc.spmm(absA, np.abs(X64)), 2e-6)
    _close(out.double().cpu().numpy(), orc.sddmm(A, G64, X64), orc.sddmm(A, np.abs(G64), np.abs(X64)), 2e-6)
    assert torch.all(yb[:offset] == 0) and torch.all(yb[offset + A.n_rows * k:] == 0)


def _hub_columns():
    """The transpose of _edge(): columns of 2600 and 4100 entries, so A^T has rows past the
    256-entry limit (the segment kernels), besides empty and short ones."""
    A = _edge(5)
    trp, tc, tv, _ = orc.transpose(A)
    return Csr(A.n_cols, A.n_rows, trp.astype(np.int32), tc.astype(np.int32), tv)


@pytest.mark.parametrize("k", [8, 64, "max"])
@pytest.mark.parametrize("vt,gt,xt", [("f64", torch.float64, torch.float64), ("f32", torch.float32, torch.float32),
                                      ("f32", torch.float32, torch.bfloat16), ("f32", torch.bfloat16, torch.bfloat16),
                                      ("bf16", torch.float32, torch.bfloat16)])
@pytest.mark.parametrize("name", ["hub", "rmat"])
def test_spmm_sddmm_transposed(mats, name, vt, gt, xt, k, dev):
    """One pass over A^T for both gradients: dX equals the separate product on A^T bit for
    bit (the same sums in the same order), dvals (in A's entry order) equals the oracle's
    SDDMM and the separate SDDMM on A within the accumulation bound."""
    A = _hub_columns() if name == "hub" else mats[name]
    V = 16 // torch.tensor([], dtype=gt).element_size()
    k = 64 * V if k == "max" else k
    rng = np.random.default_rng(k)
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[vt]
    vals = torch.tensor(rng.standard_normal(A.nnz), device=dev).to(tdt)
    G = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).to(gt)
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(xt)
    M = S.csr(A, vt, spmv_layout=1)
    M.set_values(vals)
    AT = M.transpose()
    AT.set_values(vals)
    dX = torch.full_like(X, float("nan"))
    dv = torch.full((A.nnz,), float("nan"), dtype=tdt, device=dev)
    assert AT.spmm_sddmm_transposed(G, X, dX, dv)
    dX2 = torch.full_like(X, float("nan"))
    AT.spmm(G, dX2)
    dv2 = torch.zeros(A.nnz, dtype=tdt, device=dev)
    M.sddmm(G, X, dv2)
    torch.cuda.synchronize()
    assert torch.equal(dX, dX2)
    Gn, Xn = G.double().cpu().numpy(), X.double().cpu().numpy()
    tol = 1e-13 if gt == torch.float64 else 2e-6
    if vt == "bf16":
        tol += 2 ** -8                                  # the BF16 output rounds each value once
    scale = orc.sddmm(A, np.abs(Gn), np.abs(Xn))
    _close(dv.double().cpu().numpy(), orc.sddmm(A, Gn, Xn), scale, tol)
    _close(dv.double().cpu().numpy(), dv2.double().cpu().numpy(), scale, 2 * tol)


@pytest.mark.parametrize("k", ["odd", "wide"])
def test_spmm_sddmm_transposed_declines(mats, k, dev):
    """k not a multiple of G's 16-byte vector, or wider than 64 vectors: the call returns False
    and writes nothing; a handle that is not a transpose is refused."""
    A = mats["rmat"]
    k = 6 if k == "odd" else 260
    rng = np.random.default_rng(3)
    M = S.csr(A, "f32", spmv_layout=1)
    AT = M.transpose()
    G = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).float()
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).float()
    dX = torch.full_like(X, 7.0)
    dv = torch.full((A.nnz,), 7.0, device=dev)
    assert not AT.spmm_sddmm_transposed(G, X, dX, dv)
    torch.cuda.synchronize()
    assert torch.all(dX == 7.0) and torch.all(dv == 7.0)
    with pytest.raises(Exception):
        M.spmm_sddmm_transposed(X, G, dX, dv)        # A is square: the shapes pass, the handle does not


def test_layer_backward_fused_matches_separate(dev):
    """The layer's backward (one fused pass) gives the gradients of the separate calls."""
    A = rmat_symnorm(11, 16, seed=3)
    pat = AG.SparsePattern(A, "f32")
    rng = np.random.default_rng(4)
    vals = torch.tensor(A.val, device=dev, requires_grad=True)
    X = torch.tensor(rng.standard_normal((A.n_cols, 64)), device=dev).to(torch.bfloat16).requires_grad_()
    G = torch.tensor(rng.standard_normal((A.n_rows, 64)), device=dev).float()
    Y = AG.spmm(pat, vals, X, out_dtype=torch.float32)
    Y.backward(G)
    AT = pat.AT
    AT.set_values(vals.detach())
    dX = torch.empty_like(X)
    AT.spmm(G, dX)
    dv = torch.empty(A.nnz, device=dev)
    pat.A.sddmm(G, X.detach(), dv)
    torch.cuda.synchronize()
    assert torch.equal(X.grad, dX)
    Gn, Xn = G.double().cpu().numpy(), X.detach().double().cpu().numpy()
    A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
    _close(vals.grad.double().cpu().numpy(), dv.double().cpu().numpy(), orc.sddmm(A64, np.abs(Gn), np.abs(Xn)), 4e-6)


@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["hub", "rmat"])
def test_transposed_values_after_fused_backward(mats, name, vt, dev):
    """After a fused backward the transposed handle keeps its permutation in two-pass form:
    later set_values calls use it and must give the values the plain gather gives (bit for
    bit, seen through the product), and later fused calls stay correct."""
    A = _hub_columns() if name == "hub" else mats[name]
    k = 16
    rng = np.random.default_rng(11)
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[vt]
    gt = torch.float64 if vt == "f64" else torch.float32
    M = S.csr(A, vt, spmv_layout=1)
    AT = M.transpose()
    G = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).to(gt)
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(gt)
    dX = torch.empty_like(X)
    dv = torch.empty(A.nnz, dtype=tdt, device=dev)
    assert AT.spmm_sddmm_transposed(G, X, dX, dv)     # builds the two-pass plans
    for step in range(3):
        vals = torch.tensor(rng.standard_normal(A.nnz), device=dev).to(tdt)
        AT.set_values(vals)                            # the plan's path
        fresh = S.csr(A, vt, spmv_layout=1).transpose()
        fresh.set_values(vals)                         # the gather's path
        Y1, Y2 = torch.empty_like(X), torch.empty_like(X)
        AT.spmm(G, Y1)
        fresh.spmm(G, Y2)
        dv.fill_(float("nan"))
        assert AT.spmm_sddmm_transposed(G, X, dX, dv)
        torch.cuda.synchronize()
        assert torch.equal(Y1, Y2)
        assert torch.equal(dX, Y2)
        Gn, Xn = G.double().cpu().numpy(), X.double().cpu().numpy()
        tol = (1e-13 if vt == "f64" else 2e-6) + (2 ** -8 if vt == "bf16" else 0)
        _close(dv.double().cpu().numpy(), orc.sddmm(A, Gn, Xn), orc.sddmm(A, np.abs(Gn), np.abs(Xn)), tol)
        fresh.close()


def test_fused_backward_graph_replays(mats, dev):
    """After its first call (which builds the handle's permutations), the fused backward can be
    captured: three replays, with eager work and synchronisations between them, each equal to
    the eager call bit for bit."""
    A = _hub_columns()
    k = 32
    rng = np.random.default_rng(21)
    M = S.csr(A, "f32", spmv_layout=1)
    AT = M.transpose()
    vals = torch.tensor(rng.standard_normal(A.nnz), device=dev).float()
    AT.set_values(vals)
    G = torch.tensor(rng.standard_normal((A.n_rows, k)), device=dev).float()
    X = torch.tensor(rng.standard_normal((A.n_cols, k)), device=dev).to(torch.bfloat16)
    dX0, dv0 = torch.empty_like(X), torch.empty(A.nnz, device=dev)
    assert AT.spmm_sddmm_transposed(G, X, dX0, dv0)     # eager: builds the plans
    torch.cuda.synchronize()
    dX, dv = torch.empty_like(X), torch.empty(A.nnz, device=dev)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            AT.set_values(vals, stream=s.cuda_stream)
            AT.spmm_sddmm_transposed(G, X, dX, dv, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        dX.fill_(float("nan"))
        dv.fill_(float("nan"))
        junk = torch.randn(1000, device=dev) * 2.0          # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(dX, dX0), rep
        assert torch.equal(dv, dv0), rep
        del junk
