"""SpMV parity: HIP library (through the C ABI) vs the CPU oracle and the reference layouts.

Tolerances (relative to the row condition scale s_r = sum_j |a_rj x_j|, computed in
extended precision by the oracle):
  FP64 values, FP64 accumulation:            |y - y*| <= 1e-13 * s_r
  FP32 values, FP32 (compensated) sums:      |y - y*(rounded inputs)| <= 4e-7 * s_r
  BF16 values, FP32 (compensated) sums:      |y - y*(rounded inputs)| <= 4e-7 * s_r
where y*(rounded inputs) is the exact product of the values and x as actually stored.
"""
import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd import spmv as S
from sparseops_amd.gen import Csr, laplacian_2d, powerlaw, uniform_x

pytestmark = pytest.mark.gpu

ALPHA, BETA = 0.8, 0.7


def _y0(n):
    return 0.25 * ((np.arange(n) * 7919) % 13) - 1.0


def _bf16_round(v):
    """Round-to-nearest-even to bf16 (8 significant bits), computed from the float64 value."""
    m, e = np.frexp(np.asarray(v, dtype=np.float64))
    return np.ldexp(np.rint(np.ldexp(m, 8)), e - 8)


def _run(M, x, y0, dev, alpha=ALPHA, beta=BETA):
    dt = M.xy_dtype
    xt = torch.tensor(x, dtype=dt, device=dev)
    yt = torch.tensor(y0, dtype=dt, device=dev)
    M.spmv(xt, yt, alpha, beta)
    torch.cuda.synchronize()
    return yt.cpu().double().numpy()


def _check(A, M, x, y0, vt, alpha=ALPHA, beta=BETA, dev=None):
    y = _run(M, x, y0, dev, alpha, beta)
    if vt == "f64":
        ref, s = orc.spmv_exact(A, x, alpha, beta, y0)
        tol = 1e-13
    else:
        vals = A.val.astype(np.float32).astype(np.float64) if vt == "f32" else _bf16_round(A.val)
        xr = x.astype(np.float32).astype(np.float64)
        y0r = y0.astype(np.float32).astype(np.float64)
        ref, s = orc.spmv_exact(A, xr, alpha, beta, y0r, values=vals)
        tol = 4e-7
    scale = np.abs(alpha) * s + np.abs(beta) * np.abs(y0) + 1e-300
    err = np.abs(y - ref) / scale
    assert np.all(err <= tol), f"max conditioned error {err.max():.3e} > {tol}"
    return y


def _edge_matrix(seed=5):
    """Rows of every kernel path: empty, short, stream-block, medium (>512), long (>2048)."""
    rng = np.random.default_rng(seed)
    n = 6000
    lens = rng.integers(0, 40, n)
    lens[::97] = 0
    lens[10] = 700
    lens[11] = 3000
    lens[3000] = 513
    lens[4000] = 2049
    lens[5999] = 5000
    rows, cols = [], []
    for r, L in enumerate(lens):
        c = np.sort(rng.choice(n, size=L, replace=False))
        rows.append(np.full(L, r))
        cols.append(c)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    return Csr(n, n, rp.astype(np.int32), cols.astype(np.int32), rng.standard_normal(rows.size))


@pytest.fixture(scope="module")
def mats():
    return {
        "lap": laplacian_2d(64),
        "pl": powerlaw(20000, lmax=3000, seed=11),
        "edge": _edge_matrix(),
    }


def test_laplacian_closed_form(dev):
    A = laplacian_2d(256)
    M = S.csr(A, "f64")
    y = _run(M, np.ones(A.n_cols), np.zeros(A.n_rows), dev, 1.0, 0.0)
    deg = np.diff(A.rowptr) - 1            # neighbours per row
    np.testing.assert_array_equal(y, 4.0 - deg)   # A*1 is exact in FP64


@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["lap", "pl", "edge"])
def test_csr(mats, name, vt, dev):
    """CSR row kernels (the swept layout held off)."""
    A = mats[name]
    x = uniform_x(A.n_cols, 1)
    with S.spmv_layout("sliced"):
        M = S.csr(A, vt)
    assert M.sweep is None
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)


@pytest.mark.parametrize("wgs,maxrows", [(1, 20480), (256, 20480), (40, 777)])
@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["pl", "edge", "wide", "lap"])
def test_csr_swept(mats, name, vt, wgs, maxrows, dev):
    """CSR handles with the swept SpMV layout beside the CSR arrays: SpMV on the swept layout,
    new values through set_values reach both copies (padding slots stay zero), and SpMM keeps
    using the CSR arrays."""
    A = _sweep_matrix(mats, name)
    with S.spmv_layout("swept"), S.sweep_workgroups(wgs), S.sweep_max_rows(maxrows):
        M = S.csr(A, vt)
    assert M.sweep is not None and M.info[13] == M.sweep[3]
    x = uniform_x(A.n_cols, 8)
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    y = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))
    vals2 = np.random.default_rng(5).standard_normal(A.nnz)
    M.set_values(torch.tensor(vals2, device=dev))
    A2 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, vals2)
    _check(A2, M, x, _y0(A.n_rows), vt, dev=dev)
    if vt == "f64":
        X = np.random.default_rng(6).standard_normal((A.n_cols, 3))
        Y = torch.zeros((A.n_rows, 3), dtype=torch.float64, device=dev)
        M.spmm(torch.tensor(X, device=dev), Y)
        torch.cuda.synchronize()
        for j in range(3):
            ref, sc = orc.spmv_exact(A2, X[:, j])
            assert np.all(np.abs(Y[:, j].cpu().numpy() - ref) <= 1e-13 * sc + 1e-300)


@pytest.mark.parametrize("vt", ["f64", "f32"])
def test_csr_swept_set_values_ordering(mats, vt, dev):
    """set_values refreshes the swept copy on its own stream: (1) an SpMV captured into a graph
    before set_values and replayed after it sees the new values; (2) two streams that both wait
    for set_values and then each run an SpMV both see them.  Each replay is checked against the
    oracle, with eager work and host synchronisations between the replays."""
    A = mats["pl"]
    with S.spmv_layout("swept"), S.sweep_workgroups(64):
        M = S.csr(A, vt)
    assert M.sweep is not None and M.sweep[1] > 1
    dt = M.xy_dtype
    x = uniform_x(A.n_cols, 21)
    xt = torch.tensor(x, dtype=dt, device=dev)
    yt = torch.zeros(A.n_rows, dtype=dt, device=dev)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        M.spmv(xt, yt, 1.0, 0.0)          # warm-up outside the capture
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            M.spmv(xt, yt, 1.0, 0.0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(22)
    for rep in range(3):
        vals = rng.standard_normal(A.nnz)
        M.set_values(torch.tensor(vals, device=dev))
        _ = xt.abs().sum().item()
        yt.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        A2 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, vals)
        if vt == "f64":
            ref, sc = orc.spmv_exact(A2, x)
            tol = 1e-13
        else:
            ref, sc = orc.spmv_exact(A2, x.astype(np.float32).astype(np.float64),
                                     values=vals.astype(np.float32).astype(np.float64))
            tol = 4e-7
        err = np.abs(yt.cpu().double().numpy() - ref) / (sc + 1e-300)
        assert np.all(err <= tol), f"replay {rep}: conditioned error {err.max():.3e}"
    # two streams ordered after set_values' stream by events
    vals = rng.standard_normal(A.nnz)
    s0, s1, s2 = torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream()
    y1 = torch.full((A.n_rows,), float("nan"), dtype=dt, device=dev)
    y2 = torch.full((A.n_rows,), float("nan"), dtype=dt, device=dev)
    vt_dev = torch.tensor(vals, device=dev)
    torch.cuda.synchronize()
    M.set_values(vt_dev, stream=s0.cuda_stream)
    ev = torch.cuda.Event()
    ev.record(s0)
    s1.wait_event(ev)
    s2.wait_event(ev)
    # distinct partial planes: run the two SpMVs on separate handles' planes by serialising
    # the second after the first (calls on one handle share its planes and must be ordered)
    M.spmv(xt, y1, 1.0, 0.0, stream=s1.cuda_stream)
    ev1 = torch.cuda.Event()
    ev1.record(s1)
    s2.wait_event(ev1)
    M.spmv(xt, y2, 1.0, 0.0, stream=s2.cuda_stream)
    torch.cuda.synchronize()
    A2 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, vals)
    xr = x if vt == "f64" else x.astype(np.float32).astype(np.float64)
    vr = vals if vt == "f64" else vals.astype(np.float32).astype(np.float64)
    ref, sc = orc.spmv_exact(A2, xr, values=vr)
    tol = 1e-13 if vt == "f64" else 4e-7
    for y in (y1, y2):
        err = np.abs(y.cpu().double().numpy() - ref) / (sc + 1e-300)
        assert np.all(err <= tol), f"stream-ordered SpMV: conditioned error {err.max():.3e}"


def test_csr_create_layout(mats, dev):
    """spops_csr_create_layout: 1 builds no swept copy (SpMM / SDDMM handles), 2 always does,
    whatever the process-wide layout setting."""
    A = mats["pl"]
    with S.spmv_layout("swept"):
        R = S.csr(A, "f32", spmv_layout=1)
    with S.spmv_layout("sliced"):
        W = S.csr(A, "f32", spmv_layout=2)
    assert R.sweep is None and W.sweep is not None
    x = uniform_x(A.n_cols, 3)
    _check(A, R, x, _y0(A.n_rows), "f32", dev=dev)
    _check(A, W, x, _y0(A.n_rows), "f32", dev=dev)


def test_csr_swept_auto_selection(dev):
    """CSR handles add the swept layout under the same rule as SELL (random columns over an x
    larger than the slice budget) and keep the row kernels for a band."""
    n = 1 << 16
    cols = np.sort((np.arange(n)[:, None] + np.arange(-3, 4)[None, :]) % n, axis=1)
    band = Csr(n, n, (np.arange(n + 1) * 7).astype(np.int32), cols.reshape(-1).astype(np.int32),
               np.random.default_rng(3).standard_normal(n * 7))
    P = powerlaw(n, lmax=500, seed=2)
    with S.slice_bytes(64 << 10):
        B = S.csr(band, "f32")
        R = S.csr(P, "f32")
    assert B.sweep is None and R.sweep is not None
    _check(band, B, uniform_x(n, 9), _y0(n), "f32", dev=dev)
    _check(P, R, uniform_x(n, 9), _y0(n), "f32", dev=dev)


@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["lap", "pl", "edge"])
@pytest.mark.parametrize("sigma,C", [(1024, 32), (64, 8), (16, 4), (256, 64), (48, 12)])
def test_sell(mats, name, vt, sigma, C, dev):
    A = mats[name]
    x = uniform_x(A.n_cols, 2)
    M = S.csr_to_sell_c_sigma(A, sigma, C, vt)
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    # layout parity with the oracle (pinned to the reference's csr_to_sell_c_sigma)
    align = 4 if vt == "f64" else 8
    reorder, cl, _, _, _ = orc.sell_layout(A, sigma, C, align)
    np.testing.assert_array_equal(M.layout(0), reorder)
    np.testing.assert_array_equal(M.layout(1), cl)


@pytest.mark.parametrize("vt", ["f64", "f32"])
def test_sell_c_R(mats, vt, dev):
    A = mats["pl"]
    M = S.csr_to_sell_c_R(A, 32, vt)
    _check(A, M, uniform_x(A.n_cols, 3), _y0(A.n_rows), vt, dev=dev)
    np.testing.assert_array_equal(M.layout(0), orc.sell_layout(A, M.info[11], 32, 4 if vt == "f64" else 8)[0])


@pytest.mark.parametrize("vt", ["f64", "f32"])
def test_sell_from_reference_chunks(mats, vt, dev):
    """The drop-in path: device SELL built from the reference's own host layout."""
    A = mats["edge"]
    C, sigma = 8, 64
    reorder, cl, co, scol, sval = orc.sell_layout(A, sigma, C, 4)
    ccols = [scol[co[c]:co[c] + cl[c] * C].astype(np.int32) for c in range(len(cl))]
    cvals = [sval[co[c]:co[c] + cl[c] * C] for c in range(len(cl))]
    if vt == "f32":
        cvals = [v.astype(np.float32) for v in cvals]
    M = S.sell_from_chunks(A.n_rows, A.n_cols, A.nnz, C, reorder.astype(np.int32), cl.astype(np.int32), ccols, cvals, vt)
    _check(A, M, uniform_x(A.n_cols, 4), _y0(A.n_rows), vt, dev=dev)


@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["lap", "edge"])
def test_ell(mats, name, vt, dev):
    A = mats[name]
    _check(A, S.csr_to_ell(A, vt), uniform_x(A.n_cols, 5), _y0(A.n_rows), vt, dev=dev)


@pytest.mark.parametrize("vt", ["f64", "f32"])
@pytest.mark.parametrize("bR,bC", [(16, 4), (16, 8), (4, 4), (1, 1), (8, 8), (3, 5)])
def test_bsr(mats, vt, bR, bC, dev):
    A = mats["edge"]
    _check(A, S.csr_to_bsr(A, bR, bC, vt), uniform_x(A.n_cols, 6), _y0(A.n_rows), vt, dev=dev)


def test_beta_zero_does_not_read_y(mats, dev):
    A = mats["pl"]
    M = S.csr(A, "f64")
    x = uniform_x(A.n_cols, 7)
    y0 = np.full(A.n_rows, np.nan)
    y = _run(M, x, y0, dev, 1.0, 0.0)
    ref, _ = orc.spmv_exact(A, x)
    assert np.all(np.isfinite(y))
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_local_columns_stay_unsliced(dev):
    """A banded matrix: each eighth of the rows touches a narrow band of x, so the layout
    keeps one slice even when x is larger than the per-slice budget."""
    n = 1 << 16
    cols = np.sort((np.arange(n)[:, None] + np.arange(-3, 4)[None, :]) % n, axis=1)
    A = Csr(n, n, (np.arange(n + 1) * 7).astype(np.int32), cols.reshape(-1).astype(np.int32),
            np.random.default_rng(3).standard_normal(n * 7))
    with S.slice_bytes(64 << 10), S.spmv_layout("sliced"):     # x is 256 KiB in FP32
        M = S.csr_to_sell_c_sigma(A, 1024, 32, "f32")
        R = S.csr_to_sell_c_sigma(powerlaw(n, lmax=500, seed=2), 1024, 32, "f32")
    assert M.slices == 1 and R.slices == 4
    _check(A, M, uniform_x(n, 9), _y0(n), "f32", dev=dev)


def test_empty_and_tiny(dev):
    for n in (1, 2, 63, 65):
        A = Csr(n, n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
        for M in (S.csr(A), S.csr_to_sell_c_sigma(A, 32, 32), S.csr_to_ell(A), S.csr_to_bsr(A, 4, 4)):
            y = _run(M, np.ones(n), np.ones(n), dev, 1.0, 0.5)
            np.testing.assert_array_equal(y, 0.5)


def _skewed_matrix(seed=9):
    """Half of every row's entries in columns [0, 64), the rest spread over 40000 columns:
    the equal-entry column cuts of the sliced layout crowd into the first 64 columns."""
    rng = np.random.default_rng(seed)
    n_rows, n_cols = 5000, 40000
    lens = rng.integers(0, 24, n_rows)
    lens[7], lens[8], lens[4999] = 2500, 300, 64
    cols = []
    for L in lens:
        h = min(L // 2, 64)
        a = rng.choice(64, size=h, replace=False)
        b = 64 + rng.choice(n_cols - 64, size=L - h, replace=False)
        cols.append(np.sort(np.concatenate([a, b])))
    rp = np.zeros(n_rows + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    cols = np.concatenate(cols)
    return Csr(n_rows, n_cols, rp.astype(np.int32), cols.astype(np.int32), rng.standard_normal(cols.size))


@pytest.mark.parametrize("P", [1, 2, 4, 8])
@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["pl", "edge", "skew"])
@pytest.mark.parametrize("fmt", ["sell", "ell"])
def test_column_slices(mats, name, vt, P, fmt, dev):
    """The XCD-sliced SELL/ELL layout at every slice count, including the partial-sum pass."""
    A = mats[name] if name != "skew" else _skewed_matrix()
    # the slice count follows the x bytes (in 128-byte lines) one eighth of the rows touches;
    # these matrices have random columns, so an eighth of the rows touches all of x
    xb = -(-A.n_cols * (8 if vt == "f64" else 4) // 128) * 128
    with S.slice_bytes(-(-xb // P)), S.spmv_layout("sliced"):
        M = S.csr_to_sell_c_sigma(A, 256, 32, vt) if fmt == "sell" else S.csr_to_ell(A, vt)
    assert M.slices == P
    x = uniform_x(A.n_cols, 8)
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    # beta == 0 must not read y, also through the partial-sum pass
    y = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))
    # repeated calls reuse the partial buffer: same result bit for bit
    y2 = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    np.testing.assert_array_equal(y, y2)


def _wide_matrix(seed=13):
    """20000 x 600000 with random columns: two row blocks of the swept layout, three column
    blocks (2^18 columns each at 2^14 rows per block), rows of 0 to 3000 entries."""
    rng = np.random.default_rng(seed)
    n_rows, n_cols = 20000, 600000
    lens = rng.integers(0, 16, n_rows)
    lens[::101] = 0
    lens[5], lens[16384], lens[19999] = 3000, 700, 2049
    cols = np.concatenate([np.sort(rng.choice(n_cols, size=L, replace=False)) for L in lens])
    rp = np.zeros(n_rows + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    return Csr(n_rows, n_cols, rp.astype(np.int32), cols.astype(np.int32), rng.standard_normal(cols.size))


def _sweep_matrix(mats, name):
    return _wide_matrix() if name == "wide" else _skewed_matrix() if name == "skew" else mats[name]


@pytest.mark.parametrize("wgs,maxrows", [(1, 20480), (256, 20480), (40, 777)])
@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["pl", "edge", "skew", "wide", "lap"])
@pytest.mark.parametrize("fmt", ["sell", "ell"])
def test_swept(mats, name, vt, fmt, wgs, maxrows, dev):
    """The swept layout (row blocks summed in LDS, entries sorted by column): one workgroup
    per row block (wgs=1), blocks shared by up to 16 workgroups whose partial planes a second
    pass adds (wgs=256), and many short blocks (at most 777 rows, about 40 workgroups)."""
    A = _sweep_matrix(mats, name)
    with S.spmv_layout("swept"), S.sweep_workgroups(wgs), S.sweep_max_rows(maxrows):
        M = S.csr_to_sell_c_sigma(A, 256, 32, vt) if fmt == "sell" else S.csr_to_ell(A, vt)
    rbits, H, n_rb, groups = M.sweep
    bmin = -(-A.n_rows // maxrows)
    assert M.slices == 0 and groups >= -(-A.nnz // 512) and n_rb >= bmin and (1 << rbits) >= -(-A.n_rows // n_rb)
    assert H == min(16, max(1, wgs // bmin))
    x = uniform_x(A.n_cols, 8)
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    # beta == 0 never reads y; the per-block counters reset, so repeated calls agree
    y = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))
    for _ in range(2):
        _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    if fmt == "sell":   # the reference layout arrays do not depend on the device layout
        align = 4 if vt == "f64" else 8
        reorder, cl, _, _, _ = orc.sell_layout(A, 256, 32, align)
        np.testing.assert_array_equal(M.layout(0), reorder)
        np.testing.assert_array_equal(M.layout(1), cl)


@pytest.mark.parametrize("T,wgs,maxrows", [(4, 256, 20480), (12, 64, 1000), (64, 7, 777), (100000, 256, 3000)])
@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["pl", "edge", "wide"])
@pytest.mark.parametrize("fmt", ["sell", "ell", "csr"])
def test_swept_two_class(mats, name, vt, fmt, T, wgs, maxrows, dev):
    """The two-class swept layout: rows of at most T entries and the longer rows in separate
    row blocks (row map from positions to rows), blocks of one share writing y and blocks of
    several shares writing planes that the gathering pass adds; T above every row length puts
    all rows in one class of row-limit blocks.  alpha/beta, beta == 0 (y not read), repeats."""
    A = _sweep_matrix(mats, name)
    with S.spmv_layout("swept"), S.sweep_workgroups(wgs), S.sweep_max_rows(maxrows), S.sweep_split(T):
        M = (S.csr(A, vt) if fmt == "csr" else
             S.csr_to_sell_c_sigma(A, 256, 32, vt) if fmt == "sell" else S.csr_to_ell(A, vt))
    split, nwg, planes, tasks = M.sweep_classes
    rbits, H, n_rb, groups = M.sweep
    assert split == T and nwg >= n_rb and n_rb >= -(-A.n_rows // maxrows) and groups >= -(-A.nnz // 512)
    assert nwg <= max(wgs, n_rb) and (H > 1) == (planes > 0) and (tasks > 0) == (planes > 0)
    x = uniform_x(A.n_cols, 8)
    for _ in range(2):
        _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    y = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))


@pytest.mark.parametrize("n", [20001, 20002, 20003, 20004])
@pytest.mark.parametrize("vt", ["f32", "bf16", "f64"])
def test_swept_reduce_row_counts(n, vt, dev):
    """The pass adding the partial planes: four rows per thread for FP32 planes when n % 4 == 0,
    one row per thread otherwise (and for FP64); alpha, beta and beta == 0 (y not read)."""
    A = powerlaw(n, lmax=400, seed=n)
    with S.spmv_layout("swept"), S.sweep_workgroups(256):
        M = S.csr_to_sell_c_sigma(A, 256, 32, vt)
    assert M.sweep[1] > 1
    x = uniform_x(A.n_cols, 12)
    _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    y = _run(M, x, np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))


def test_swept_empty_tiny_and_nonfinite(dev):
    """Empty and one-row matrices; x holding inf/nan outside the gathered columns must not
    reach y (padding slots never gather)."""
    with S.spmv_layout("swept"):
        for n in (1, 2, 63, 65):
            A = Csr(n, n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
            M = S.csr_to_sell_c_sigma(A, 32, 32)
            y = _run(M, np.ones(n), np.ones(n), dev, 1.0, 0.5)
            np.testing.assert_array_equal(y, 0.5)
        n = 3000
        rng = np.random.default_rng(4)
        cols = np.stack([np.sort(1 + rng.choice(n - 1, size=5, replace=False)) for _ in range(n)])
        rp = np.arange(n + 1, dtype=np.int32) * 5
        A = Csr(n, n, rp, cols.reshape(-1).astype(np.int32), rng.standard_normal(n * 5))
        for vt in ("f64", "f32"):
            M = S.csr_to_sell_c_sigma(A, 64, 32, vt)
            x = uniform_x(n, 3)
            x[0] = np.inf                        # column 0 holds no entry
            ref, _ = orc.spmv_exact(A, np.where(np.isfinite(x), x, 0.0))
            y = _run(M, x, np.zeros(n), dev, 1.0, 0.0)
            assert np.all(np.isfinite(y))
            np.testing.assert_allclose(y, ref, rtol=0, atol=(1e-12 if vt == "f64" else 1e-5) * np.abs(ref).max())


def test_swept_auto_selection(dev):
    """The default picks the swept layout for random columns over an x larger than the slice
    budget and keeps the sliced layout for a band."""
    n = 1 << 16
    cols = np.sort((np.arange(n)[:, None] + np.arange(-3, 4)[None, :]) % n, axis=1)
    band = Csr(n, n, (np.arange(n + 1) * 7).astype(np.int32), cols.reshape(-1).astype(np.int32),
               np.random.default_rng(3).standard_normal(n * 7))
    with S.slice_bytes(64 << 10):
        B = S.csr_to_sell_c_sigma(band, 1024, 32, "f32")
        R = S.csr_to_sell_c_sigma(powerlaw(n, lmax=500, seed=2), 1024, 32, "f32")
    assert B.sweep is None and B.slices == 1
    assert R.sweep is not None and R.slices == 0
    _check(band, B, uniform_x(n, 9), _y0(n), "f32", dev=dev)


@pytest.mark.parametrize("vt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("bR,bC", [(8, 8), (16, 16), (16, 8), (8, 12), (24, 16), (40, 8), (32, 32), (64, 64)])
@pytest.mark.parametrize("name", ["edge", "pl"])
def test_bsr_mfma(mats, name, vt, bR, bC, dev):
    """Dense blocks of at least 8 x 8 run on the matrix cores (v_mfma_f64_16x16x4f64); partial
    16-row tiles, several column groups, blocks past the VALU kernel's 256 entries, and the last
    block column past n_cols (x there must read 0)."""
    A = mats[name]
    with S.bsr_mfma("always"):
        M = S.csr_to_bsr(A, bR, bC, vt)
    assert M.info[9] == 1                       # the MFMA kernel runs it
    _check(A, M, uniform_x(A.n_cols, 6), _y0(A.n_rows), vt, dev=dev)
    y = _run(M, uniform_x(A.n_cols, 6), np.full(A.n_rows, np.nan), dev, 1.0, 0.0)
    assert np.all(np.isfinite(y))


@pytest.mark.parametrize("vt", ["f64", "f32"])
@pytest.mark.parametrize("bR,bC", [(8, 8), (16, 16)])
def test_bsr_mfma_matches_valu(mats, vt, bR, bC, dev):
    """The same blocks on the VALU kernel and on the matrix cores agree within the sum
    tolerance, and the switch keeps shapes the VALU kernel cannot hold on the matrix cores."""
    A = mats["pl"]
    x = uniform_x(A.n_cols, 12)
    V = S.csr_to_bsr(A, bR, bC, vt)                 # default: blocks the VALU kernel holds stay there
    with S.bsr_mfma("always"):
        M = S.csr_to_bsr(A, bR, bC, vt)
    assert V.info[9] == 0 and M.info[9] == 1
    assert S.csr_to_bsr(A, 32, 16, vt).info[9] == 1  # 512 entries: past the VALU kernel, on the matrix cores
    yv = _check(A, V, x, _y0(A.n_rows), vt, dev=dev)
    ym = _check(A, M, x, _y0(A.n_rows), vt, dev=dev)
    ref, s = orc.spmv_exact(A, x, ALPHA, BETA, _y0(A.n_rows))
    tol = 2e-13 if vt == "f64" else 8e-7
    assert np.all(np.abs(yv - ym) <= tol * (ALPHA * s + BETA * np.abs(_y0(A.n_rows)) + 1e-300))
    with S.bsr_mfma("never"), pytest.raises(RuntimeError):
        S.csr_to_bsr(A, 32, 32, vt)                 # 1024 entries: too large for the VALU kernel
