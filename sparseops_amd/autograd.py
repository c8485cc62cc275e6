"""Differentiable SpMM Y = A(vals) X with hand-written backward kernels (torch.autograd).

The pattern of A is fixed and lives on the GPU; its values are a tensor that may require
gradients.  Forward and backward call the C ABI of libsparseops_hip.so:
  Y     = A X                      spops_spmm_values on A (vals read in place)
  dX    = A^T G                    spops_spmm_typed on a cached transposed handle whose values
                                   follow vals through the entry permutation
  dvals = SDDMM(G, X) on A's pattern   (dvals[e] = <G[row_e, :], X[col_e, :]>, spops_sddmm)
When both gradients are needed they come from one pass over the transposed handle
(spops_spmm_sddmm_transposed), where k and the alignment allow it.
Accumulation is FP64 when the values or X are FP64, otherwise FP32 (BF16 X / G are widened).
The transpose semantics are the reference's csr_transpose (include/sparse_conversion.h:2067-2108)
without its duplicate merging: the pattern has no duplicates.
"""
from __future__ import annotations

import torch

from . import spmv as S
from .gen import Csr


class SparsePattern:
    """A CSR pattern on the device with values of `value_type` (f64, f32 or bf16 storage)
    and its transpose, built once."""

    def __init__(self, A: Csr, value_type="f32"):
        # SpMM, SDDMM and the transpose read the CSR arrays only: no swept SpMV copy to refresh
        self.A = S.csr(A, value_type, spmv_layout=1)
        self.AT = self.A.transpose()
        self.n_rows, self.n_cols, self.nnz = A.n_rows, A.n_cols, A.nnz
        self.value_dtype = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[value_type]

    def close(self):
        self.A.close()
        self.AT.close()


class _SpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals: torch.Tensor, X: torch.Tensor, pat: SparsePattern, out_dtype):
        X = X.contiguous()
        Y = torch.empty(pat.n_rows, X.shape[1], dtype=out_dtype or X.dtype, device=X.device)
        v = vals.detach()
        if v.dtype == pat.value_dtype and v.is_contiguous():
            pat.A.spmm_values(v, X, Y)          # the values read in place, no copy
        else:
            pat.A.set_values(v)
            pat.A.spmm(X, Y)
        ctx.save_for_backward(vals, X)
        ctx.pat = pat
        return Y

    @staticmethod
    def backward(ctx, G: torch.Tensor):
        vals, X = ctx.saved_tensors
        pat = ctx.pat
        G = G.contiguous()
        dvals = dX = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1] and vals.dtype == pat.value_dtype:
            # both gradients in one pass over A^T: each gathered row of G serves dX and dvals
            pat.AT.set_values(vals.detach())
            dX = torch.empty_like(X)
            dvals = torch.empty(pat.nnz, dtype=vals.dtype, device=vals.device)
            if pat.AT.spmm_sddmm_transposed(G, X, dX, dvals):
                return dvals, dX, None, None
            pat.AT.spmm(G, dX)
            pat.A.sddmm(G, X, dvals)
            return dvals, dX, None, None
        if ctx.needs_input_grad[1]:
            # one permuting gather of the values, then the product on contiguous values (reading
            # them through the permutation inside the gather-bound product measured slower)
            pat.AT.set_values(vals.detach())
            dX = torch.empty_like(X)
            pat.AT.spmm(G, dX)
        if ctx.needs_input_grad[0]:
            dvals = torch.empty(pat.nnz, dtype=vals.dtype, device=vals.device)
            pat.A.sddmm(G, X, dvals)
        return dvals, dX, None, None


def spmm(pat: SparsePattern, vals: torch.Tensor, X: torch.Tensor, out_dtype=None) -> torch.Tensor:
    """Y = A(vals) X, differentiable in vals and X.  out_dtype defaults to X's dtype
    (pass torch.float32 to keep FP32 outputs for BF16 X)."""
    if vals.numel() != pat.nnz:
        raise ValueError("vals must have nnz elements")
    return _SpMM.apply(vals, X, pat, out_dtype)
