"""Matrix I/O: MatrixMarket (.mtx) and a lossless binary CSR container.

MatrixMarket reading follows the reference's reader semantics
(utils/sparse_io.cpp:26-169): 1-based indices become 0-based, ``pattern`` entries
read as 1.0, ``symmetric`` files mirror every off-diagonal entry, and entries keep
file order within a row (the counting sort of include/sparse_conversion.h:38-99).
Duplicates are kept unless ``compact`` is set (include/sparse_io.h:52).

Writing uses ``%.17g`` so values round-trip exactly (the reference's writer uses
``%20.16g``, utils/mmio.cpp:383, which does not).

Binary CSR ("SPOCSR01"): int64 n_rows, n_cols, nnz; int64 rowptr[n_rows+1];
int64 col[nnz]; float64 val[nnz]; little-endian.
"""
from __future__ import annotations

import os

import numpy as np

from .gen import Csr

MAGIC = b"SPOCSR01"


def write_csr_bin(path, A: Csr) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([A.n_rows, A.n_cols, A.nnz], dtype=np.int64).tofile(f)
        A.rowptr.astype(np.int64).tofile(f)
        A.col.astype(np.int64).tofile(f)
        A.val.astype(np.float64).tofile(f)


def read_csr_bin(path, index_dtype=np.int64) -> Csr:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a SPOCSR01 file")
        n_rows, n_cols, nnz = np.fromfile(f, dtype=np.int64, count=3)
        rowptr = np.fromfile(f, dtype=np.int64, count=n_rows + 1)
        col = np.fromfile(f, dtype=np.int64, count=nnz)
        val = np.fromfile(f, dtype=np.float64, count=nnz)
    return Csr(int(n_rows), int(n_cols), rowptr.astype(index_dtype), col.astype(index_dtype), val)


def write_mtx(path, A: Csr) -> None:
    """General real coordinate file, entries in CSR order, exact %.17g values."""
    rows = np.repeat(np.arange(A.n_rows, dtype=np.int64), A.row_lengths()) + 1
    cols = A.col.astype(np.int64) + 1
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.n_rows} {A.n_cols} {A.nnz}\n")
        np.savetxt(f, np.column_stack([rows, cols, A.val.astype(np.float64)]),
                   fmt=["%d", "%d", "%.17g"])


def read_mtx(path, index_dtype=np.int64, compact: bool = False) -> Csr:
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0] != "%%MatrixMarket":
            raise ValueError(f"{path}: could not process Matrix Market banner")
        obj, fmt, field, sym = (s.lower() for s in banner[1:5])
        if obj != "matrix" or fmt != "coordinate" or field not in ("real", "integer", "pattern"):
            raise ValueError(f"{path}: only sparse real/integer/pattern coordinate matrices are supported")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n_rows, n_cols, nnz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2) if nnz else np.zeros((0, 3))
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    if sym == "symmetric":
        off = rows != cols
        # each off-diagonal entry is followed by its mirror (utils/sparse_io.cpp:124-166)
        rep = np.where(off, 2, 1)
        r2 = np.repeat(rows, rep)
        c2 = np.repeat(cols, rep)
        v2 = np.repeat(vals, rep)
        second = np.zeros(r2.size, dtype=bool)
        second[np.cumsum(rep)[off] - 1] = True
        r2[second], c2[second] = c2[second].copy(), np.repeat(rows, rep)[second]
        rows, cols, vals = r2, c2, v2
    order = np.argsort(rows, kind="stable")          # row-stable counting sort
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=n_rows)
    rowptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    A = Csr(n_rows, n_cols, rowptr, cols, vals)
    if compact:
        A = sum_duplicates(A)
    return A.astype_index(index_dtype)


def sum_duplicates(A: Csr) -> Csr:
    """compact=true reading (include/sparse_operation.h:21-73 through coo_to_csr): inside each
    row the entries of one column are added in file order, exact-zero sums are dropped, and the
    columns come out in reverse order of their first appearance (the order the reference's
    linked list of seen columns yields)."""
    nnz = A.nnz
    rows = np.repeat(np.arange(A.n_rows, dtype=np.int64), A.row_lengths())
    key = rows * max(A.n_cols, 1) + A.col.astype(np.int64)
    uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
    vals = np.zeros(uk.size)
    np.add.at(vals, inv, A.val)                     # entries of a group added in entry order
    r = uk // max(A.n_cols, 1)
    keep = vals != 0.0
    r, c, v, first = r[keep], (uk % max(A.n_cols, 1))[keep], vals[keep], first[keep]
    order = np.lexsort((-first, r))                 # by row, then latest first appearance first
    r, c, v = r[order], c[order], v[order]
    counts = np.bincount(r, minlength=A.n_rows) if nnz else np.zeros(A.n_rows, np.int64)
    rowptr = np.zeros(A.n_rows + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return Csr(A.n_rows, A.n_cols, rowptr.astype(A.rowptr.dtype), c.astype(A.col.dtype), v)


def write_clustering(out_dir: str, name: str, perm, offsets) -> tuple[str, str]:
    """A row clustering as record_permutation_lsh writes it (test/record_permutation_lsh.cpp:
    160-181): <out_dir>/<name>.perm and <name>.offsets, one integer per line."""
    os.makedirs(out_dir, exist_ok=True)
    pp, po = os.path.join(out_dir, f"{name}.perm"), os.path.join(out_dir, f"{name}.offsets")
    for path, arr in ((pp, perm), (po, offsets)):
        with open(path, "w") as f:
            f.write("".join(f"{int(v)}\n" for v in np.asarray(arr, dtype=np.int64)))
    return pp, po


def read_clustering(lsh_dir: str, name: str, n_rows: int):
    """(perm, offsets) of <lsh_dir>/<name>.perm / .offsets; the permutation padded with
    (size, size+1, ...) or cut to n_rows, as include/sparse_conversion.h:2012-2052 does (a
    Matrix Market header and size line, if present, are skipped)."""
    with open(os.path.join(lsh_dir, f"{name}.perm")) as f:
        lines = f.read().splitlines()
    if lines and "%%MatrixMarket" in lines[0]:
        i = 1
        while i < len(lines) and lines[i].startswith("%"):
            i += 1
        lines = lines[i + 1:]
    perm = np.array([int(t) for ln in lines for t in ln.split()], dtype=np.int64)
    if n_rows > 0:
        if perm.size < n_rows:
            perm = np.concatenate([perm, np.arange(perm.size, n_rows, dtype=np.int64)])
        perm = perm[:n_rows]
    with open(os.path.join(lsh_dir, f"{name}.offsets")) as f:
        offsets = np.array([int(t) for t in f.read().split()], dtype=np.int64)
    return perm, offsets
