"""Matrix features of the reference's feature extractor (MTX<I,V>, include/sparse_features.h:
13-326, src/sparse_features.cpp:170-850) computed on the device (spops_matrix_features).

    f = features(A)            # dict name -> value, names as the reference prints them
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from .gen import Csr

NAMES = [
    "num_rows", "num_cols", "real_nnz", "nnz_ratio", "num_of_diags", "nnz_closed2diag_ratio",
    "nnz_lower", "nnz_upper", "nnz_diagonal",
    "NoneZero_Row_Ratio", "min_nnz_each_row", "max_nnz_each_row", "ave_nnz_each_row", "var_nnz_each_row",
    "standard_dev_row", "P-ratio_row", "Gini_coeff_row",
    "NoneZero_Col_Ratio", "min_nnz_each_col", "max_nnz_each_col", "ave_nnz_each_col", "var_nnz_each_col",
    "standard_dev_col", "P-ratio_col", "Gini_coeff_col",
    "diagonal_dominant_ratio", "max_value_offdiag", "max_value_diagonal", "row_variability", "col_variability",
    "Number of Tiles", "tile_rows", "tile_cols",
    "NE_ratio_tiles_", "min_nnz_each_tiles", "max_nnz_each_tiles", "t_ave_nnz_tiles", "t_var_nnz_tiles",
    "t_stand_dev_tiles", "P-ratio_tiles", "Gini_nnz_tiles",
    "NE_ratio_RB", "min_nnz_each_RB", "max_nnz_each_RB", "t_ave_nnz_RB", "t_var_nnz_RB", "t_stand_dev_RB",
    "P-ratio_RB", "Gini_nnz_RB",
    "NE_ratio_CB", "min_nnz_each_CB", "max_nnz_each_CB", "t_ave_nnz_CB", "t_var_nnz_CB", "t_stand_dev_CB",
    "P-ratio_CB", "Gini_nnz_CB",
    "uniq_R", "uniq_C", "potReuseR", "potReuseC",
]
assert len(NAMES) == 61


def features_array(rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor, n_rows: int, n_cols: int) -> np.ndarray:
    """Device CSR (int32 or int64 indices, float32 or float64 values) -> 61 doubles."""
    if rowptr.dtype != col.dtype or rowptr.dtype not in (torch.int32, torch.int64):
        raise ValueError("rowptr and col must share an int32 or int64 dtype")
    if val.dtype not in (torch.float32, torch.float64):
        raise ValueError("values must be float32 or float64")
    out = np.zeros(61)
    L.call("spops_matrix_features", rowptr.data_ptr(), col.data_ptr(), val.data_ptr(),
           L.INDEX_I32 if rowptr.dtype == torch.int32 else L.INDEX_I64,
           L.VALUE_F32 if val.dtype == torch.float32 else L.VALUE_F64, int(n_rows), int(n_cols),
           out.ctypes.data, L.current_stream())
    return out


def features(A: Csr, device="cuda:0") -> dict:
    rp = torch.as_tensor(np.ascontiguousarray(A.rowptr, dtype=np.int64), device=device)
    ci = torch.as_tensor(np.ascontiguousarray(A.col, dtype=np.int64), device=device)
    v = torch.as_tensor(np.ascontiguousarray(A.val, dtype=np.float64), device=device)
    return dict(zip(NAMES, features_array(rp, ci, v, A.n_rows, A.n_cols)))
