"""Batched SpMM over many FP32 matrices in one persistent launch (optionally replayed from a
captured HIP graph), with the per-matrix format / tile autotuner of the library.

    B = BatchSpMM(mats, ks, tune=True)        # mats: list of Csr, ks: RHS count per matrix
    X = B.new_x(); Y = B.new_y()              # packed device buffers
    B.x_view(X, i)                            # (n_cols_i, k_i) view of matrix i's block (row stride ld_i)
    B.spmm(X, Y)                              # Y_i = A_i X_i for every i, one launch
    B.capture(X, Y, stream); B.replay(stream) # the same launch as a HIP graph
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from .gen import Csr

FORMATS = {0: "csr_groups", 1: "chunks"}


class BatchSpMM:
    def __init__(self, mats: list[Csr], ks: list[int], tune: bool = True):
        n = len(mats)
        if len(ks) != n:
            raise ValueError("one k per matrix")
        self.n = n
        self.shapes = [(A.n_rows, A.n_cols) for A in mats]
        self.ks = [int(k) for k in ks]
        keep = []
        rps, cis, vs = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
        for i, A in enumerate(mats):
            if A.nnz >= 2 ** 31 - 1 or max(A.n_rows, A.n_cols) >= 2 ** 31 - 1:
                raise ValueError(f"matrix {i}: rows, columns and entries must stay below 2^31")
            rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
            ci = np.ascontiguousarray(A.col, dtype=np.int32)
            v = np.ascontiguousarray(A.val, dtype=np.float32)
            keep += [rp, ci, v]
            rps[i], cis[i], vs[i] = rp.ctypes.data, ci.ctypes.data, v.ctypes.data
        nr = np.array([A.n_rows for A in mats], np.int64)
        nc = np.array([A.n_cols for A in mats], np.int64)
        kk = np.array(self.ks, np.int64)
        h = ctypes.c_void_p()
        L.call("spops_batch_create", ctypes.byref(h), n, nr.ctypes.data, nc.ctypes.data, kk.ctypes.data, L.INDEX_I32,
               rps, cis, vs, L.VALUE_F32, 1 if tune else 0)
        self._h = h
        vals = [ctypes.c_int64() for _ in range(5)]
        L.call("spops_batch_info", h, *[ctypes.byref(v) for v in vals])
        self.x_len, self.y_len, self.flop, self.algo_bytes, self.ntasks = [v.value for v in vals]
        self.x_off = np.zeros(n, np.int64)
        self.y_off = np.zeros(n, np.int64)
        self.ld = np.zeros(n, np.int32)
        self.fmt = np.zeros(n, np.int32)
        self.lanes = np.zeros(n, np.int32)
        self.sigma = np.zeros(n, np.int32)
        self.cls = np.zeros(n, np.int32)
        L.call("spops_batch_plan", h, self.x_off.ctypes.data, self.y_off.ctypes.data, self.ld.ctypes.data,
               self.fmt.ctypes.data, self.lanes.ctypes.data, self.sigma.ctypes.data, self.cls.ctypes.data)

    def new_x(self, device="cuda:0"):
        return torch.zeros(max(self.x_len, 1), dtype=torch.float32, device=device)

    def new_y(self, device="cuda:0"):
        return torch.zeros(max(self.y_len, 1), dtype=torch.float32, device=device)

    def x_view(self, X: torch.Tensor, i: int) -> torch.Tensor:
        o, ld, n = int(self.x_off[i]), int(self.ld[i]), self.shapes[i][1]
        return X[o:o + n * ld].view(n, ld)[:, :self.ks[i]]

    def y_view(self, Y: torch.Tensor, i: int) -> torch.Tensor:
        o, ld, n = int(self.y_off[i]), int(self.ld[i]), self.shapes[i][0]
        return Y[o:o + n * ld].view(n, ld)[:, :self.ks[i]]

    def _check(self, X, Y):
        for T, n in ((X, self.x_len), (Y, self.y_len)):
            if T.dtype != torch.float32 or not T.is_cuda or not T.is_contiguous() or T.numel() < n \
                    or T.data_ptr() % 16:
                raise ValueError("X / Y must be contiguous, 16-byte aligned float32 device buffers of x_len / y_len "
                                 "elements")

    def spmm(self, X: torch.Tensor, Y: torch.Tensor, alpha: float = 1.0, beta: float = 0.0, stream=None):
        self._check(X, Y)
        s = L.current_stream() if stream is None else stream
        L.call("spops_batch_spmm", self._h, float(alpha), X.data_ptr(), float(beta), Y.data_ptr(), s)
        return Y

    def capture(self, X: torch.Tensor, Y: torch.Tensor, stream, alpha: float = 1.0, beta: float = 0.0):
        self._check(X, Y)
        L.call("spops_batch_capture", self._h, float(alpha), X.data_ptr(), float(beta), Y.data_ptr(), stream)

    def replay(self, stream):
        L.call("spops_batch_replay", self._h, stream)

    def formats(self):
        return [FORMATS[int(f)] for f in self.fmt]

    def close(self):
        if self._h:
            L.lib.spops_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
