"""spops_spmm_typed_noalloc (the SpMM entry point that allocates nothing: its long-row
partial sums live in a static device array) against spops_spmm_typed, through the C ABI.
The two run the same kernels in the same order, so the results must be bit-identical."""
import ctypes

import numpy as np
import pytest
import torch

from sparseops_amd import _lib as L
from sparseops_amd import spmv as S
from sparseops_amd.gen import Csr, powerlaw

pytestmark = pytest.mark.gpu

TYPES = {torch.float64: L.VALUE_F64, torch.float32: L.VALUE_F32, torch.bfloat16: L.VALUE_BF16}


def _spmm(fn, M, X, Y, alpha=1.0, beta=0.0):
    k = X.shape[1]
    return getattr(L.lib, fn)(M._h, k, ctypes.c_double(alpha), ctypes.c_void_p(X.data_ptr()), TYPES[X.dtype],
                     ctypes.c_double(beta), ctypes.c_void_p(Y.data_ptr()), TYPES[Y.dtype],
                     ctypes.c_void_p(L.current_stream()))


def _long_rows(n, seed):
    """Rows of every SpMM path, with many rows longer than 256 entries (segments)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, n)
    lens[::37] = rng.integers(257, 3000, len(lens[::37]))
    cols = [np.sort(rng.choice(n, size=min(int(L_), n), replace=False)) for L_ in lens]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum([len(c) for c in cols], out=rp[1:])
    col = np.concatenate(cols)
    return Csr(n, n, rp.astype(np.int32), col.astype(np.int32), rng.standard_normal(col.size))


@pytest.mark.parametrize("vt,xt,yt", [("f64", torch.float64, torch.float64), ("f32", torch.bfloat16, torch.float32),
                                      ("f32", torch.float32, torch.float32)])
def test_noalloc_matches_spmm_typed_twice(vt, xt, yt, dev):
    """Two calls in a row, the second on a smaller matrix with fewer columns of X; each equals
    the allocating entry point's output bit for bit (NaN-filled outputs first)."""
    for n, k, seed, (a, b) in ((6000, 64, 1, (1.0, 0.0)), (2500, 8, 2, (0.5, -1.25))):
        A = _long_rows(n, seed)
        M = S.csr(A, vt, spmv_layout=1)
        g = torch.Generator(device="cpu").manual_seed(seed)
        X = torch.randn(n, k, generator=g).to(xt).to(dev)
        Y0 = torch.randn(n, k, generator=g).to(yt).to(dev)
        want = Y0.clone()
        assert _spmm("spops_spmm_typed", M, X, want, a, b) == 0
        got = Y0.clone()
        assert _spmm("spops_spmm_typed_noalloc", M, X, got, a, b) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"n={n} k={k}: noalloc differs"
        M.close()


def test_noalloc_rejects_beyond_its_scratch(dev):
    """One row of 2^20 entries is 1024 segments: k = 16384 FP64 sums need 134 MB > 64 MiB (the call is
    refused before it touches X or Y)."""
    n = 1 << 20
    A = Csr(1, n, np.array([0, n], np.int64), np.arange(n, dtype=np.int64), np.ones(n))
    M = S.csr(A, "f64", spmv_layout=1)
    X = torch.zeros(16, dtype=torch.float64, device=dev)
    Y = torch.zeros(16, dtype=torch.float64, device=dev)
    st = L.lib.spops_spmm_typed_noalloc(M._h, 16384, ctypes.c_double(1.0), ctypes.c_void_p(X.data_ptr()), L.VALUE_F64,
                                        ctypes.c_double(0.0), ctypes.c_void_p(Y.data_ptr()), L.VALUE_F64,
                                        ctypes.c_void_p(L.current_stream()))
    assert st == 5 and b"static scratch" in L.lib.spops_last_error()
    M.close()


def test_noalloc_graph_replays(dev):
    """Captured into a graph (it allocates nothing), replayed three times with eager work and
    synchronisations between the replays; every replay equals the eager call."""
    A = powerlaw(20000, lmax=3000, seed=11)
    M = S.csr(A, "f32", spmv_layout=1)
    g = torch.Generator(device="cpu").manual_seed(3)
    X = torch.randn(A.n_cols, 32, generator=g).to(dev)
    Y = torch.full((A.n_rows, 32), float("nan"), device=dev)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        assert _spmm("spops_spmm_typed_noalloc", M, X, Y) == 0     # first call outside the capture
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        _spmm("spops_spmm_typed_noalloc", M, X, Y)
    for rep in range(3):
        X.copy_(torch.randn(A.n_cols, 32, generator=g).to(dev))
        Y.fill_(float("nan"))
        _ = X.abs().sum().item()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        want = torch.empty_like(Y)
        assert _spmm("spops_spmm_typed", M, X, want) == 0
        torch.cuda.synchronize()
        assert torch.equal(Y, want), f"replay {rep} differs"
    M.close()
