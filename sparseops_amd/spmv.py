"""Python handle over the C-ABI SpMV matrices (device-resident CSR / SELL-C-sigma / ELL / BSR).

Mirrors the reference's operator surface for the SpMV path:
``csr_to_sell_c_sigma`` / ``csr_to_sell_c_R`` / ``csr_to_ell`` / ``csr_to_bsr`` build a
device matrix (include/sparse_conversion.h:195, :345, :491, :777) and
``SpOps(alpha, A, x, beta, y)`` computes y = alpha*A*x + beta*y
(src/spmv_csr.cpp:165, spmv_ell.cpp:225, spmv_sell_c_sigma.cpp:190,
spmv_sell_c_R.cpp:190, spmv_bsr.cpp:215) on device tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from .gen import Csr

VALUE_TYPES = {"f64": L.VALUE_F64, "f32": L.VALUE_F32, "bf16": L.VALUE_BF16}
XY_DTYPE = {L.VALUE_F64: torch.float64, L.VALUE_F32: torch.float32, L.VALUE_BF16: torch.float32}
FORMAT_NAMES = {L.FORMAT_CSR: "csr", L.FORMAT_SELL: "sell", L.FORMAT_ELL: "ell", L.FORMAT_BSR: "bsr"}


TORCH_TYPES = {torch.float64: L.VALUE_F64, torch.float32: L.VALUE_F32, torch.bfloat16: L.VALUE_BF16}


def _rhs_check(X: torch.Tensor, rows: int, name: str) -> int:
    if X.dim() != 2 or X.shape[0] != rows or not X.is_contiguous() or not X.is_cuda:
        raise ValueError(f"{name} must be a contiguous device tensor of shape ({rows}, k)")
    if X.dtype not in TORCH_TYPES:
        raise TypeError(f"{name} must be float64, float32 or bfloat16")
    return int(X.shape[1])


def _vt(v) -> int:
    return VALUE_TYPES[v] if isinstance(v, str) else int(v)


def _index_arrays(A: Csr):
    it = L.INDEX_I32 if A.rowptr.dtype == np.int32 else L.INDEX_I64
    dt = np.int32 if it == L.INDEX_I32 else np.int64
    rp = np.ascontiguousarray(A.rowptr, dtype=dt)
    ci = np.ascontiguousarray(A.col, dtype=dt)
    return it, rp, ci


def _host_values(A: Csr):
    v = np.ascontiguousarray(A.val)
    if v.dtype == np.float32:
        return v, L.VALUE_F32
    return np.ascontiguousarray(v, dtype=np.float64), L.VALUE_F64


class SpMatrix:
    """A device matrix handle (owns the HIP allocations made by the library)."""

    def __init__(self, handle: ctypes.c_void_p, keep=()):
        self._h = handle
        self._keep = keep
        info = (ctypes.c_int64 * 16)()
        L.call("spops_matrix_info", self._h, info)
        self.info = list(info)
        self.format = FORMAT_NAMES[self.info[0]]
        self.n_rows, self.n_cols, self.nnz = self.info[1], self.info[2], self.info[3]
        self.value_type = self.info[4]
        self.xy_dtype = XY_DTYPE[self.value_type]

    @property
    def handle(self):
        return self._h

    @property
    def device_bytes(self) -> int:
        return self.info[5]

    @property
    def algo_bytes(self) -> int:
        """Compulsory bytes of one SpMV (beta = 0): indices, values, x once, y once."""
        return self.info[6]

    def spmv(self, x: torch.Tensor, y: torch.Tensor, alpha: float = 1.0, beta: float = 0.0, stream=None):
        if x.dtype != self.xy_dtype or y.dtype != self.xy_dtype:
            raise TypeError(f"x/y must be {self.xy_dtype} for this matrix")
        if not (x.is_cuda and y.is_cuda and x.is_contiguous() and y.is_contiguous()):
            raise ValueError("x and y must be contiguous device tensors")
        if x.numel() < self.n_cols or y.numel() < self.n_rows:
            raise ValueError("x or y too short")
        s = L.current_stream() if stream is None else stream
        L.call("spops_spmv", self._h, float(alpha), x.data_ptr(), float(beta), y.data_ptr(), s)
        return y

    # ---- multi-RHS and differentiable forms (CSR) ----------------------------------------
    def spmm(self, X: torch.Tensor, Y: torch.Tensor, alpha: float = 1.0, beta: float = 0.0, stream=None):
        """Y = alpha*A*X + beta*Y; X (n_cols x k) and Y (n_rows x k) row-major device tensors
        in FP64, FP32 or BF16 (independently)."""
        k = _rhs_check(X, self.n_cols, "X")
        if Y.dim() != 2 or Y.shape[0] != self.n_rows or Y.shape[1] != k or not Y.is_contiguous() or not Y.is_cuda:
            raise ValueError("Y must be a contiguous device tensor of shape (n_rows, k)")
        s = L.current_stream() if stream is None else stream
        L.call("spops_spmm_typed", self._h, k, float(alpha), X.data_ptr(), TORCH_TYPES[X.dtype], float(beta),
               Y.data_ptr(), TORCH_TYPES[Y.dtype], s)
        return Y

    def spmm_values(self, vals: torch.Tensor, X: torch.Tensor, Y: torch.Tensor, alpha: float = 1.0,
                    beta: float = 0.0, stream=None):
        """spmm with the matrix values read from `vals` (nnz elements of the stored value type)
        instead of the stored ones, which stay unchanged; for a transpose() handle, vals are in
        the original matrix's entry order (as for set_values)."""
        k = _rhs_check(X, self.n_cols, "X")
        if Y.dim() != 2 or Y.shape[0] != self.n_rows or Y.shape[1] != k or not Y.is_contiguous() or not Y.is_cuda:
            raise ValueError("Y must be a contiguous device tensor of shape (n_rows, k)")
        if (vals.numel() != self.nnz or not vals.is_contiguous() or not vals.is_cuda
                or TORCH_TYPES.get(vals.dtype) != self.info[4]):
            raise ValueError("vals must be a contiguous device tensor of nnz elements of the stored value type")
        s = L.current_stream() if stream is None else stream
        L.call("spops_spmm_values", self._h, vals.data_ptr(), k, float(alpha), X.data_ptr(), TORCH_TYPES[X.dtype],
               float(beta), Y.data_ptr(), TORCH_TYPES[Y.dtype], s)
        return Y

    def sddmm(self, G: torch.Tensor, X: torch.Tensor, out: torch.Tensor, stream=None):
        """out[e] = <G[row_e, :], X[col_e, :]> over A's pattern (the gradient of the values)."""
        k = _rhs_check(X, self.n_cols, "X")
        if G.dim() != 2 or G.shape != (self.n_rows, k) or not G.is_contiguous() or not G.is_cuda:
            raise ValueError("G must be a contiguous device tensor of shape (n_rows, k)")
        if out.numel() != self.nnz or not out.is_contiguous() or not out.is_cuda:
            raise ValueError("out must be a contiguous device tensor of nnz elements")
        s = L.current_stream() if stream is None else stream
        L.call("spops_sddmm", self._h, k, G.data_ptr(), TORCH_TYPES[G.dtype], X.data_ptr(), TORCH_TYPES[X.dtype],
               out.data_ptr(), TORCH_TYPES[out.dtype], s)
        return out

    def spmm_sddmm_transposed(self, G: torch.Tensor, X: torch.Tensor, dX: torch.Tensor, dvals: torch.Tensor,
                              stream=None) -> bool:
        """On a transpose() handle AT of A: dX = AT G and dvals[e] = <G[row_e, :], X[col_e, :]> over
        A's entries (A's entry order, this handle's value type) in one pass.  Returns False, having
        done nothing, where the fused kernels do not apply (k or alignment); then spmm on this
        handle and sddmm on A compute the same."""
        k = _rhs_check(G, self.n_cols, "G")
        if X.dim() != 2 or X.shape != (self.n_rows, k) or not X.is_contiguous() or not X.is_cuda:
            raise ValueError("X must be a contiguous device tensor of shape (n_rows, k)")
        if dX.shape != X.shape or dX.dtype != X.dtype or not dX.is_contiguous() or not dX.is_cuda:
            raise ValueError("dX must be a contiguous device tensor shaped and typed like X")
        if (dvals.numel() != self.nnz or not dvals.is_contiguous() or not dvals.is_cuda
                or TORCH_TYPES.get(dvals.dtype) != self.info[4]):
            raise ValueError("dvals must be a contiguous device tensor of nnz elements of the stored value type")
        if X.dtype not in TORCH_TYPES:
            raise TypeError("X must be float64, float32 or bfloat16")
        s = L.current_stream() if stream is None else stream
        st = L.lib.spops_spmm_sddmm_transposed(self._h, k, G.data_ptr(), TORCH_TYPES[G.dtype], X.data_ptr(),
                                               TORCH_TYPES[X.dtype], dX.data_ptr(), dvals.data_ptr(), s)
        if st == L.NOT_SUPPORTED:
            return False
        L.check(st, "spops_spmm_sddmm_transposed")
        return True

    def transpose(self) -> "SpMatrix":
        """A^T as a CSR handle whose values follow this matrix's (see set_values)."""
        h = ctypes.c_void_p()
        L.call("spops_csr_transpose", self._h, ctypes.byref(h))
        return SpMatrix(h)

    def set_values(self, vals: torch.Tensor, stream=None):
        """Overwrite the stored values (cast to the stored type) from a device tensor of nnz
        elements; for a transpose() handle, vals are in the original matrix's entry order."""
        if vals.numel() != self.nnz or not vals.is_contiguous() or not vals.is_cuda:
            raise ValueError("vals must be a contiguous device tensor of nnz elements")
        s = L.current_stream() if stream is None else stream
        L.call("spops_csr_set_values", self._h, vals.data_ptr(), TORCH_TYPES[vals.dtype], s)

    @property
    def slices(self) -> int:
        """Column slices of the SELL/ELL device layout (0 for the swept layout)."""
        return self.info[12] if self.format in ("sell", "ell") else 1

    @property
    def sweep(self):
        """Swept layout parameters (row bits, workgroups per row block, row blocks, entry
        groups), or None when the matrix's SpMV uses another layout (CSR handles keep it beside
        the CSR arrays, which SpMM, SDDMM and the transpose use)."""
        if self.format not in ("sell", "ell", "csr"):
            return None
        out = np.zeros(4, dtype=np.int64)
        L.call("spops_matrix_layout", self._h, 2, out.ctypes.data)
        return tuple(int(v) for v in out) if out[1] > 0 else None

    @property
    def sweep_classes(self):
        """Row classes of the swept layout (split length T, 0 for one class; workgroups;
        partial-plane elements; tasks of the plane-adding pass), or None without one."""
        if self.sweep is None:
            return None
        out = np.zeros(4, dtype=np.int64)
        L.call("spops_matrix_layout", self._h, 3, out.ctypes.data)
        return tuple(int(v) for v in out)

    def layout(self, which: int) -> np.ndarray:
        n = self.n_rows if which == 0 else self.info[14]
        out = np.zeros(n, dtype=np.int32)
        L.call("spops_matrix_layout", self._h, which, out.ctypes.data)
        return out

    def close(self):
        if self._h:
            L.lib.spops_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _create(fn: str, A: Csr, value_type, *extra) -> SpMatrix:
    it, rp, ci = _index_arrays(A)
    v, in_vt = _host_values(A)
    h = ctypes.c_void_p()
    L.call(fn, ctypes.byref(h), A.n_rows, A.n_cols, A.nnz, it, rp.ctypes.data, ci.ctypes.data,
           v.ctypes.data, in_vt, _vt(value_type), *extra)
    return SpMatrix(h)


def csr(A: Csr, value_type="f64", spmv_layout=None) -> SpMatrix:
    """CSR handle.  spmv_layout None follows the process-wide tuning; 0 chooses from the matrix,
    1 keeps the row kernels only (handles used for SpMM / SDDMM / the transpose), 2 adds the
    swept SpMV layout."""
    if spmv_layout is None:
        return _create("spops_csr_create", A, value_type)
    return _create("spops_csr_create_layout", A, value_type, int(spmv_layout))


def csr_to_sell_c_sigma(A: Csr, sigma: int, C: int, value_type="f64") -> SpMatrix:
    return _create("spops_sell_create", A, value_type, int(sigma), int(C))


def csr_to_sell_c_R(A: Csr, C: int, value_type="f64") -> SpMatrix:
    """SELL-C-R: one slice over the whole matrix (sigma = C * ceil(n / C) >= n)."""
    sigma = int(C) * max(1, -(-A.n_rows // int(C)))
    return _create("spops_sell_create", A, value_type, sigma, int(C))


def csr_to_ell(A: Csr, value_type="f64") -> SpMatrix:
    return _create("spops_ell_create", A, value_type)


def csr_to_bsr(A: Csr, bR: int, bC: int, value_type="f64") -> SpMatrix:
    return _create("spops_bsr_create", A, value_type, int(bR), int(bC))


def sell_from_chunks(n_rows, n_cols, nnz, C, reorder, chunk_len, chunk_cols, chunk_vals, value_type="f64") -> SpMatrix:
    """Device SELL from the reference's own host layout (per-chunk row-major arrays)."""
    reorder = np.ascontiguousarray(reorder)
    it = L.INDEX_I32 if reorder.dtype == np.int32 else L.INDEX_I64
    dt = reorder.dtype
    chunk_len = np.ascontiguousarray(chunk_len, dtype=dt)
    cols = [np.ascontiguousarray(c, dtype=dt) for c in chunk_cols]
    vals = [np.ascontiguousarray(v) for v in chunk_vals]
    in_vt = L.VALUE_F32 if vals and vals[0].dtype == np.float32 else L.VALUE_F64
    vals = [np.ascontiguousarray(v, dtype=np.float32 if in_vt == L.VALUE_F32 else np.float64) for v in vals]
    nch = len(cols)
    cp = (ctypes.c_void_p * nch)(*[c.ctypes.data for c in cols])
    vp = (ctypes.c_void_p * nch)(*[v.ctypes.data for v in vals])
    h = ctypes.c_void_p()
    L.call("spops_sell_create_from_chunks", ctypes.byref(h), n_rows, n_cols, nnz, C, nch, it,
           reorder.ctypes.data, chunk_len.ctypes.data, cp, vp, in_vt, _vt(value_type))
    return SpMatrix(h)


class tuning:
    """Context manager (or plain call) setting one process-wide tuning value of the library
    (spops_set_tuning) for matrices created while it is active; the old value comes back on
    exit."""

    def __init__(self, key: int, value: int):
        old = ctypes.c_int64(0)
        L.call("spops_set_tuning", key, int(value), ctypes.byref(old))
        self.key, self.old = key, old.value

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        L.call("spops_set_tuning", self.key, self.old, None)
        return False


def slice_bytes(nbytes: int) -> tuning:
    """Bytes of x per column slice of the sliced SELL/ELL layout (default 2 MiB, one XCD's
    share of L2)."""
    return tuning(L.TUNE_SLICE_BYTES, nbytes)


LAYOUTS = {"auto": 0, "sliced": 1, "swept": 2}


def spmv_layout(kind: str) -> tuning:
    """Device layout of SELL/ELL matrices created inside: "auto" (per matrix), "sliced"
    (column-sliced stream) or "swept" (row blocks summed in LDS, entries of a block sorted by
    column).  CSR matrices keep their row kernels under "sliced", add the swept layout for
    SpMV under "swept", and choose as SELL does under "auto"."""
    return tuning(L.TUNE_SPMV_LAYOUT, LAYOUTS[kind])


def sweep_max_rows(n: int) -> tuning:
    """Rows of the largest row block of the swept layout (1..20480, default 20480)."""
    return tuning(L.TUNE_SWEEP_MAX_ROWS, n)


BSR_MFMA = {"never": 0, "large": 1, "always": 2}


def bsr_mfma(mode: str) -> tuning:
    """Which BSR matrices with dense blocks of at least 8 x 8 run on the matrix cores: "never",
    "large" (default: blocks past the VALU kernel's 256 entries) or "always"."""
    return tuning(L.TUNE_BSR_MFMA, BSR_MFMA[mode])


def sweep_workgroups(n: int) -> tuning:
    """Workgroups a swept-layout SpMV aims for (default 256); row blocks are shared by up to 16."""
    return tuning(L.TUNE_SWEEP_WORKGROUPS, n)


def sweep_split(T: int) -> tuning:
    """Row blocks of the swept layout: 0 (default) consecutive rows of equal entry counts,
    T > 0 the rows of at most T entries and the longer rows in separate blocks."""
    return tuning(L.TUNE_SWEEP_SPLIT, T)


def SpOps(alpha: float, A: SpMatrix, x: torch.Tensor, beta: float, y: torch.Tensor):
    """y = alpha*A*x + beta*y (the reference's SpOps_<fmt> argument order)."""
    return A.spmv(x, y, alpha, beta)
