"""Seeded synthetic workloads for the SpMV / LeSpGEMM hot path.

Every generator is deterministic for a given seed (numpy PCG64 streams) and
returns a :class:`Csr` whose rows hold unique, ascending column indices, which
is what the reference's MatrixMarket reader produces for a ``general`` file
written in row order (utils/sparse_io.cpp:26-169, include/sparse_conversion.h:38-99
keeps file order within a row).

Workloads (sizes are the ones the throughput figures are quoted on):

* ``laplacian_2d(256)``   5-point Laplacian, n = 65,536, nnz = 326,656.
* ``powerlaw(2**21)``     n = 2,097,152, row lengths P(L) ~ L^-2.1 on [3, 8192],
                          columns uniform without replacement, values N(0,1).
* ``rmat(20, 16)``        R-MAT (a,b,c,d) = (.57,.19,.19,.05), duplicates merged,
                          values U(0,1].
* ``band_batch(...)``     512 band + random off-band matrices, FP32 N(0,1).
* ``rmat_symnorm(20,32)`` R-MAT with D^-1/2 A D^-1/2 values (row-degree D).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    n_rows: int
    n_cols: int
    rowptr: np.ndarray   # (n_rows+1,) int32 or int64
    col: np.ndarray      # (nnz,) same dtype as rowptr
    val: np.ndarray      # (nnz,) float64 (or float32)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def astype_index(self, dt) -> "Csr":
        return Csr(self.n_rows, self.n_cols, self.rowptr.astype(dt), self.col.astype(dt), self.val)

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowptr.astype(np.int64))


def _finish(n_rows, n_cols, rows, cols, vals, index_dtype):
    """Sort COO by (row, col) and build CSR; rows/cols must already be unique pairs."""
    key = rows.astype(np.int64) * np.int64(n_cols) + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    rows = rows[order]
    cols = cols[order]
    vals = vals[order] if vals is not None else None
    counts = np.bincount(rows, minlength=n_rows).astype(np.int64)
    rowptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return Csr(n_rows, n_cols, rowptr.astype(index_dtype), cols.astype(index_dtype), vals)


def laplacian_2d(nx: int = 256, index_dtype=np.int32) -> Csr:
    """5-point Laplacian on an nx-by-nx grid: diagonal 4, neighbours -1."""
    n = nx * nx
    idx = np.arange(n, dtype=np.int64)
    i, j = idx // nx, idx % nx
    rows = [idx]
    cols = [idx]
    vals = [np.full(n, 4.0)]
    for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        ok = (i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < nx)
        rows.append(idx[ok])
        cols.append((i[ok] + di) * nx + (j[ok] + dj))
        vals.append(np.full(int(ok.sum()), -1.0))
    return _finish(n, n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), index_dtype)


def uniform_x(n: int, seed: int = 0, dtype=np.float64) -> np.ndarray:
    """x ~ U[-1, 1) from the PCG64 stream `seed` (numpy default_rng)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(dtype)


def _unique_cols_per_row(rng, n_rows, n_cols, lens):
    """Uniform columns without replacement per row (rejection of duplicates)."""
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), lens)
    cols = rng.integers(0, n_cols, rows.size, dtype=np.int64)
    while True:
        key = rows * n_cols + cols
        order = np.argsort(key, kind="stable")
        rows, cols = rows[order], cols[order]
        dup = np.zeros(rows.size, dtype=bool)
        dup[1:] = (rows[1:] == rows[:-1]) & (cols[1:] == cols[:-1])
        nd = int(dup.sum())
        if nd == 0:
            return rows, cols
        cols[dup] = rng.integers(0, n_cols, nd, dtype=np.int64)


def powerlaw(n: int = 1 << 21, alpha: float = 2.1, lmin: int = 3, lmax: int = 8192,
             seed: int = 0, index_dtype=np.int32) -> Csr:
    """Square matrix with truncated discrete power-law row lengths, N(0,1) values."""
    rng = np.random.default_rng(seed)
    lmax = min(lmax, n)
    support = np.arange(lmin, lmax + 1, dtype=np.float64)
    p = support ** (-alpha)
    cdf = np.cumsum(p / p.sum())
    cdf[-1] = 1.0
    lens = (lmin + np.searchsorted(cdf, rng.random(n), side="right")).astype(np.int64)
    lens = np.minimum(lens, lmax)
    rows, cols = _unique_cols_per_row(rng, n, n, lens)
    vals = rng.standard_normal(rows.size)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    return Csr(n, n, rowptr.astype(index_dtype), cols.astype(index_dtype), vals)


def block_sparse(n: int = 1 << 18, b: int = 16, blocks_per_row: int = 8, seed: int = 0) -> Csr:
    """Square matrix of dense b x b blocks: every block row holds blocks_per_row distinct block
    columns, uniform without replacement; N(0,1) values; rows sorted by column."""
    rng = np.random.default_rng(seed)
    nb = -(-n // b)
    bcols = np.stack([np.sort(rng.choice(nb, size=min(blocks_per_row, nb), replace=False)) for _ in range(nb)])
    k = bcols.shape[1]
    cols = (bcols[:, None, :, None] * b + np.arange(b)[None, None, None, :]).repeat(b, axis=1)
    cols = cols.reshape(nb * b, k * b)[:n]
    keep = cols < n
    lens = keep.sum(axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    col = cols[keep].astype(np.int32)
    return Csr(n, n, rowptr.astype(np.int32), col, rng.standard_normal(col.size))


def rmat(scale: int = 20, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, d=0.05,
         seed: int = 0, index_dtype=np.int64) -> Csr:
    """R-MAT graph, duplicate edges merged, self loops kept, values U(0,1]."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    m = edge_factor * n
    rows = np.zeros(m, dtype=np.int64)
    cols = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    for lvl in range(scale):
        u = rng.random(m)
        bit = np.int64(1) << np.int64(scale - 1 - lvl)
        rbit = u >= ab                      # quadrants c, d
        cbit = ((u >= a) & (u < ab)) | (u >= abc)   # quadrants b, d
        rows += rbit * bit
        cols += cbit * bit
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    vals = 1.0 - rng.random(rows.size)
    return _finish(n, n, rows, cols, vals, index_dtype)


def random_rect(n_rows: int, n_cols: int, per_row_max: int, seed: int = 0, index_dtype=np.int64) -> Csr:
    """Rectangular matrix: U{0..per_row_max} distinct uniform columns per row, values
    N(0,1) times 10^U(-3,3) (magnitudes on both sides of 1), every tenth row with an
    explicit zero."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, per_row_max + 1, n_rows)
    lens = np.minimum(lens, n_cols)
    rows, cols = _unique_cols_per_row(rng, n_rows, n_cols, lens)
    vals = rng.standard_normal(rows.size) * 10.0 ** rng.uniform(-3, 3, rows.size)
    first = np.r_[0, np.cumsum(lens)[:-1]]
    z = first[(np.arange(n_rows) % 10 == 0) & (lens > 0)]
    vals[z] = 0.0
    return _finish(n_rows, n_cols, rows, cols, vals, index_dtype)


def transpose(A: Csr) -> Csr:
    """Exact transpose (CSC of A read as CSR); value multiset is unchanged."""
    rows = np.repeat(np.arange(A.n_rows, dtype=np.int64), A.row_lengths())
    return _finish(A.n_cols, A.n_rows, A.col.astype(np.int64), rows, A.val.copy(),
                   A.rowptr.dtype)


def rmat_symnorm(scale: int = 20, edge_factor: int = 32, seed: int = 0,
                 index_dtype=np.int32) -> Csr:
    """R-MAT pattern with values D^-1/2 A D^-1/2, D = row degree of the pattern
    (an empty row or column gets degree 1), stored as FP32."""
    A = rmat(scale, edge_factor, seed=seed, index_dtype=np.int64)
    deg = np.maximum(A.row_lengths(), 1).astype(np.float64)
    rows = np.repeat(np.arange(A.n_rows, dtype=np.int64), A.row_lengths())
    v = 1.0 / np.sqrt(deg[rows] * deg[A.col])
    return Csr(A.n_rows, A.n_cols, A.rowptr.astype(index_dtype), A.col.astype(index_dtype),
               v.astype(np.float32))


def band_matrix(n: int, half_width: int, extra_max: int, rng, dtype=np.float32,
                index_dtype=np.int32) -> Csr:
    """Band |i-j| <= half_width plus U{0..extra_max} uniform off-band entries per row
    (duplicates and extras inside the band merged).  Built without a global sort: per row,
    the sorted extras left of the band, the band, then the extras right of it."""
    w = int(half_width)
    i = np.arange(n, dtype=np.int64)
    lo = np.maximum(i - w, 0)
    blen = np.minimum(i + w, n - 1) - lo + 1
    extra = rng.integers(0, extra_max + 1, n)
    erow = np.repeat(i, extra)
    ecol = rng.integers(0, n, erow.size, dtype=np.int64)
    keep = np.abs(ecol - erow) > w
    key = np.unique(erow[keep] * n + ecol[keep])
    erow, ecol = key // n, key % n
    left = ecol < erow
    nleft = np.bincount(erow[left], minlength=n).astype(np.int64)
    nex = np.bincount(erow, minlength=n).astype(np.int64)
    counts = blen + nex
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    nnz = int(rowptr[-1])
    # band entries: position p of row r holds column p - (rowptr[r] + nleft[r]) + lo[r]
    cols = np.arange(nnz, dtype=np.int64) + np.repeat(lo - rowptr[:-1] - nleft, counts)
    # extras: rank inside the row's sorted extras
    estart = np.zeros(n + 1, np.int64)
    np.cumsum(nex, out=estart[1:])
    rank = np.arange(erow.size, dtype=np.int64) - estart[erow]
    cols[rowptr[erow] + np.where(left, rank, blen[erow] + rank)] = ecol
    vals = rng.standard_normal(nnz, dtype=np.float32 if dtype == np.float32 else np.float64).astype(dtype, copy=False)
    return Csr(n, n, rowptr.astype(index_dtype), cols.astype(index_dtype), vals)


def band_batch(count: int = 512, nmin: int = 1000, nmax: int = 64000, wmax: int = 64,
               extra_max: int = 8, kmax: int = 32, seed: int = 0):
    """Batch of band matrices: n log-uniform in [nmin, nmax], half-width U{1..wmax},
    0..extra_max random off-band nnz per row, RHS count U{1..kmax}."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    ws = rng.integers(1, wmax + 1, count)
    ks = rng.integers(1, kmax + 1, count)
    mats = [band_matrix(int(ns[t]), int(ws[t]), extra_max, rng) for t in range(count)]
    return mats, [int(k) for k in ks]
