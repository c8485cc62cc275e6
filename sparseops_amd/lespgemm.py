"""LeSpGEMM pipeline on the GPU through the C ABI (mirrors test/test_spgemm_hc_lsh.cpp:49-203).

    pairs   = lsh_candidate_pairs(A, k, bands, seed, constrain)      # MinHash + LSH (device)
    perm, offsets = hierarchical_clustering(A, pairs, cluster_size)  # host, bit-exact v0_fast / v1
    A_cl    = ClusterMatrix(A, perm, offsets)                        # csr_reorder_rows + csr_to_vlength_cluster
    C_cl    = A_cl.multiply(B)                                       # LeSpGEMM_VLength kernel 3
    C       = C_cl.to_csr(A_cl)                                      # vlength_cluster2csr + inverse permutation

    plan    = SpgemmPlan(A_cl, B)                                    # the same product in cluster panels
    for c0, c1 in plan.panel_ranges(budget_bytes):
        P = plan.numeric(c0, c1, P)                                  # clusters [c0, c1) of C_cl

    C       = spgemm_rows(dA, dB)                                    # row-wise LeSpGEMM (kernel flags 1-3)
    Cand    = topk_jaccard(dA, top_k)                                # HashSpGEMMTopK candidate pairs
    perm, offsets = hierarchical_clustering_v0(A, Cand.pairs, ...)   # map-based v0 over those pairs

Every step runs in libsparseops_hip.so; this module only moves arrays.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from .gen import Csr

c_i64 = ctypes.c_int64


def _dev(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), device=device).to(dtype)


class DeviceCsr:
    """A CSR matrix with int64 indices resident on the device (torch allocations)."""

    def __init__(self, A: Csr, device, value_dtype=torch.float64):
        self.n_rows, self.n_cols, self.nnz = A.n_rows, A.n_cols, A.nnz
        self.rowptr = _dev(A.rowptr.astype(np.int64), torch.int64, device)
        self.col = _dev(A.col.astype(np.int64), torch.int64, device)
        self.val = _dev(A.val, value_dtype, device)
        self.value_type = L.VALUE_F64 if value_dtype == torch.float64 else L.VALUE_F32


def minhash_signatures(dA: DeviceCsr, k: int = 64, seed: int = 12345) -> torch.Tensor:
    sigs = torch.empty((max(dA.n_rows, 1), k), dtype=torch.int64, device=dA.rowptr.device)
    L.call("spops_minhash_signatures", dA.rowptr.data_ptr(), dA.col.data_ptr(), L.INDEX_I64, dA.n_rows, k,
           ctypes.c_uint64(seed), sigs.data_ptr(), L.current_stream())
    return sigs[: dA.n_rows]


def lsh_candidate_pairs(dA: DeviceCsr, k: int = 64, bands: int = 16, seed: int = 12345, constrain: bool = True,
                        score_f32: bool = False):
    """(pi, pj, score) numpy arrays, sorted by (i, j); scores in float64 (or float32)."""
    h = ctypes.c_void_p()
    st = L.VALUE_F32 if score_f32 else L.VALUE_F64
    L.call("spops_lsh_candidate_pairs", dA.rowptr.data_ptr(), dA.col.data_ptr(), L.INDEX_I64, dA.n_rows, k, bands,
           ctypes.c_uint64(seed), int(constrain), st, ctypes.byref(h), L.current_stream())
    try:
        n = c_i64()
        L.call("spops_pairs_count", h, ctypes.byref(n))
        pi = np.zeros(n.value, np.int64)
        pj = np.zeros(n.value, np.int64)
        sc = np.zeros(n.value, np.float32 if score_f32 else np.float64)
        L.call("spops_pairs_download", h, pi.ctypes.data, pj.ctypes.data, sc.ctypes.data)
    finally:
        L.lib.spops_pairs_destroy(h)
    return pi, pj, sc


def hierarchical_clustering(A: Csr, pi, pj, score, cluster_size: int = 8, version: int = 0):
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.col, dtype=np.int64)
    pi = np.ascontiguousarray(pi, dtype=np.int64)
    pj = np.ascontiguousarray(pj, dtype=np.int64)
    score = np.ascontiguousarray(score)
    st = L.VALUE_F32 if score.dtype == np.float32 else L.VALUE_F64
    perm = np.zeros(A.n_rows, np.int64)
    offsets = np.zeros(A.n_rows + 1, np.int64)
    nc = c_i64()
    L.call("spops_hierarchical_clustering", version, A.n_rows, rp.ctypes.data, ci.ctypes.data, len(pi),
           pi.ctypes.data, pj.ctypes.data, score.ctypes.data, st, cluster_size, perm.ctypes.data,
           offsets.ctypes.data, ctypes.byref(nc))
    return perm, offsets[: nc.value + 1].copy()


def cluster_matching(n_rows: int, pi: torch.Tensor, pj: torch.Tensor, score: torch.Tensor, cluster_size: int = 8,
                     max_rounds: int = 256):
    """Approximate clustering on the GPU (spops_cluster_matching) of device pair arrays (int64 i, j;
    float64 or float32 scores >= 0): rounds of mutual best edges between clusters that fit
    cluster_size.  Returns (perm, offsets, rounds)."""
    for t in (pi, pj, score):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("pairs must be contiguous device tensors")
    if pi.dtype != torch.int64 or pj.dtype != torch.int64 or score.dtype not in (torch.float64, torch.float32):
        raise ValueError("pairs: int64 i, j and float64 / float32 scores")
    st = L.VALUE_F64 if score.dtype == torch.float64 else L.VALUE_F32
    perm = np.zeros(max(n_rows, 1), np.int64)
    offsets = np.zeros(n_rows + 1, np.int64)
    nc, rounds = c_i64(), c_i64()
    L.call("spops_cluster_matching", n_rows, pi.numel(), pi.data_ptr(), pj.data_ptr(), score.data_ptr(), st,
           cluster_size, max_rounds, perm.ctypes.data, offsets.ctypes.data, ctypes.byref(nc), ctypes.byref(rounds),
           L.current_stream())
    return perm[:n_rows].copy(), offsets[: nc.value + 1].copy(), rounds.value


def cluster_rows_gpu(dA: DeviceCsr, k: int = 64, bands: int = 16, seed: int = 7, constrain: bool = True,
                     cluster_size: int = 8, max_rounds: int = 256, score_f32: bool = False):
    """MinHash + LSH pairs and the approximate GPU clustering, the pairs staying on the device:
    (perm, offsets, rounds, npairs)."""
    h = ctypes.c_void_p()
    st = L.VALUE_F32 if score_f32 else L.VALUE_F64
    L.call("spops_lsh_candidate_pairs", dA.rowptr.data_ptr(), dA.col.data_ptr(), L.INDEX_I64, dA.n_rows, k, bands,
           ctypes.c_uint64(seed), int(constrain), st, ctypes.byref(h), L.current_stream())
    try:
        n = c_i64()
        L.call("spops_pairs_count", h, ctypes.byref(n))
        di, dj, ds = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        L.call("spops_pairs_device", h, ctypes.byref(di), ctypes.byref(dj), ctypes.byref(ds))
        perm = np.zeros(max(dA.n_rows, 1), np.int64)
        offsets = np.zeros(dA.n_rows + 1, np.int64)
        nc, rounds = c_i64(), c_i64()
        L.call("spops_cluster_matching", dA.n_rows, n.value, di, dj, ds, st, cluster_size, max_rounds,
               perm.ctypes.data, offsets.ctypes.data, ctypes.byref(nc), ctypes.byref(rounds), L.current_stream())
    finally:
        L.lib.spops_pairs_destroy(h)
    return perm[: dA.n_rows].copy(), offsets[: nc.value + 1].copy(), rounds.value, n.value


def hierarchical_clustering_v0(A: Csr, pi, pj, score, cluster_size: int = 8):
    """The map-based hierarchical_clustering_v0 (include/spgemm_utility.h:517-606) over pairs in
    the order of the caller's map (keys as it holds them): (perm, offsets, roots), clusters
    ordered by their root row."""
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.col, dtype=np.int64)
    pi = np.ascontiguousarray(pi, dtype=np.int64)
    pj = np.ascontiguousarray(pj, dtype=np.int64)
    score = np.ascontiguousarray(score)
    st = L.VALUE_F32 if score.dtype == np.float32 else L.VALUE_F64
    perm = np.zeros(A.n_rows, np.int64)
    offsets = np.zeros(A.n_rows + 1, np.int64)
    roots = np.zeros(max(A.n_rows, 1), np.int64)
    nc = c_i64()
    L.call("spops_hierarchical_clustering_v0", A.n_rows, rp.ctypes.data, ci.ctypes.data, len(pi), pi.ctypes.data,
           pj.ctypes.data, score.ctypes.data, st, cluster_size, perm.ctypes.data, offsets.ctypes.data,
           roots.ctypes.data, ctypes.byref(nc))
    return perm, offsets[: nc.value + 1].copy(), roots[: nc.value].copy()


class ClusterMatrix:
    """Device CSR_VlengthCluster (owned by the library)."""

    def __init__(self, handle, value_type):
        self._h = handle
        self.value_type = value_type
        self.stats = None
        self._refresh()

    def _refresh(self):
        info = (c_i64 * 8)()
        L.call("spops_cluster_info", self._h, info)
        self.rows, self.csr_rows, self.cols, self.nnzc, self.nnzv, self.max_cluster_sz = list(info)[:6]

    def checksum(self):
        """{sum of values, sum of |values|, sum of column ids, nonzero values}."""
        out = (ctypes.c_double * 4)()
        L.call("spops_cluster_checksum", self._h, out, L.current_stream())
        return {"sum": out[0], "sum_abs": out[1], "colid_sum": int(round(out[2])), "nonzeros": int(round(out[3]))}

    @classmethod
    def build(cls, dA: DeviceCsr, perm, offsets):
        perm = np.ascontiguousarray(perm, dtype=np.int64)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        h = ctypes.c_void_p()
        L.call("spops_cluster_create", ctypes.byref(h), dA.n_rows, dA.n_cols, dA.nnz, dA.rowptr.data_ptr(),
               dA.col.data_ptr(), dA.val.data_ptr(), dA.value_type, perm.ctypes.data, len(offsets) - 1,
               offsets.ctypes.data, L.current_stream())
        return cls(h, dA.value_type)

    def multiply(self, dB: DeviceCsr) -> "ClusterMatrix":
        h = ctypes.c_void_p()
        stats = (c_i64 * 16)()
        L.call("spops_lespgemm", self._h, dB.n_rows, dB.n_cols, dB.nnz, dB.rowptr.data_ptr(), dB.col.data_ptr(),
               dB.val.data_ptr(), ctypes.byref(h), stats, L.current_stream())
        C = ClusterMatrix(h, self.value_type)
        C.stats = _stats(stats)
        return C

    def download(self):
        vt = np.float64 if self.value_type == L.VALUE_F64 else np.float32
        csz = np.zeros(self.rows, np.int64)
        rp = np.zeros(self.rows + 1, np.int64)
        rpv = np.zeros(self.rows + 1, np.int64)
        col = np.zeros(self.nnzc, np.int64)
        val = np.zeros(self.nnzv, vt)
        L.call("spops_cluster_download", self._h, csz.ctypes.data, rp.ctypes.data, rpv.ctypes.data, col.ctypes.data,
               val.ctypes.data)
        return csz, rp, rpv, col, val

    def symbolic(self):
        nz, mn, mx = (np.zeros(self.rows, np.int64) for _ in range(3))
        L.call("spops_cluster_symbolic_download", self._h, nz.ctypes.data, mn.ctypes.data, mx.ctypes.data)
        return nz, mn, mx

    def to_csr(self, A_cl: "ClusterMatrix") -> Csr:
        nnz = c_i64()
        drp, dcol, dval = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        L.call("spops_cluster_to_csr", self._h, A_cl._h, ctypes.byref(nnz), ctypes.byref(drp), ctypes.byref(dcol),
               ctypes.byref(dval), L.current_stream())
        try:
            n = self.csr_rows
            vt = np.float64 if self.value_type == L.VALUE_F64 else np.float32
            rp = np.zeros(n + 1, np.int64)
            col = np.zeros(nnz.value, np.int64)
            val = np.zeros(nnz.value, vt)
            L.call("spops_memcpy_dtoh", rp.ctypes.data, drp, rp.nbytes, None)
            L.call("spops_memcpy_dtoh", col.ctypes.data, dcol, col.nbytes, None)
            L.call("spops_memcpy_dtoh", val.ctypes.data, dval, val.nbytes, None)
        finally:
            for p in (drp, dcol, dval):
                L.lib.spops_free(p)
        return Csr(n, self.cols, rp, col, val)

    def close(self):
        if self._h:
            L.lib.spops_cluster_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class dense_products:
    """Context manager: LeSpGEMM plans created inside use a dense accumulator for clusters
    with at least `products` products."""

    def __init__(self, products: int):
        self.products = int(products)

    def __enter__(self):
        old = c_i64()
        L.call("spops_set_tuning", L.TUNE_DENSE_PRODUCTS, self.products, ctypes.byref(old))
        self.old = old.value
        return self

    def __exit__(self, *exc):
        L.call("spops_set_tuning", L.TUNE_DENSE_PRODUCTS, self.old, None)


class task_products:
    """Context manager: numeric calls made inside split a large cluster's work over workgroups
    of about `products` products each (0, the library default: chosen per call, 2^15 halved
    down to 2^13 while the call has fewer than 32768 tasks)."""

    def __init__(self, products: int):
        self.products = int(products)

    def __enter__(self):
        old = c_i64()
        L.call("spops_set_tuning", L.TUNE_TASK_PRODUCTS, self.products, ctypes.byref(old))
        self.old = old.value
        return self

    def __exit__(self, *exc):
        L.call("spops_set_tuning", L.TUNE_TASK_PRODUCTS, self.old, None)


class heavy_products:
    """Context manager: plans created inside build the symbolic bitmap of clusters with at
    least `products` products with several workgroups and keep it for the emit pass (library
    default 2^20)."""

    def __init__(self, products: int):
        self.products = int(products)

    def __enter__(self):
        old = c_i64()
        L.call("spops_set_tuning", L.TUNE_HEAVY_PRODUCTS, self.products, ctypes.byref(old))
        self.old = old.value
        return self

    def __exit__(self, *exc):
        L.call("spops_set_tuning", L.TUNE_HEAVY_PRODUCTS, self.old, None)


class wide_cursors:
    """Context manager: numeric calls made inside run clusters of 1017..4096 A columns in the
    variant with 4096 LDS cursors (two workgroups per CU): mode 0 never (per-chunk searches of
    each B row in the regular kernel), 1 always, 2 (library default) when the call has at
    least 8192 such tasks."""

    def __init__(self, mode: int):
        self.mode = int(mode)

    def __enter__(self):
        old = c_i64()
        L.call("spops_set_tuning", L.TUNE_WIDE_CURSORS, self.mode, ctypes.byref(old))
        self.old = old.value
        return self

    def __exit__(self, *exc):
        L.call("spops_set_tuning", L.TUNE_WIDE_CURSORS, self.old, None)


def _stats(stats):
    s = list(stats)
    return {"b_rows_sorted_copy": s[0], "nnzc": s[1], "nnzv": s[2], "small_clusters": s[3], "large_clusters": s[4],
            "setup_ms": s[5] / 1e3, "symbolic_ms": s[6] / 1e3, "malloc_ms": s[7] / 1e3, "numeric_ms": s[8] / 1e3,
            "cluster_products": s[9], "dense_clusters": s[10], "large_tasks": s[11], "wide_tasks": s[12]}


class SpgemmPlan:
    """C = A_cl * B prepared once (B's sorted int32 copy, bins, symbolic phase of every
    cluster), then computed in cluster ranges whose output fits a memory budget."""

    def __init__(self, Acl: ClusterMatrix, dB: DeviceCsr):
        if Acl.cols != dB.n_rows:
            raise ValueError("Error: A.cols != B.rows")
        self._keep = (Acl, dB)
        self.value_type = Acl.value_type
        h = ctypes.c_void_p()
        L.call("spops_lespgemm_plan_create", Acl._h, dB.n_rows, dB.n_cols, dB.nnz, dB.rowptr.data_ptr(),
               dB.col.data_ptr(), dB.val.data_ptr(), ctypes.byref(h), L.current_stream())
        self._h = h
        info = (c_i64 * 8)()
        L.call("spops_lespgemm_plan_info", h, info)
        (self.clusters, self.nnzc, self.nnzv, self.products, self.small_clusters,
         self.large_clusters) = list(info)[:6]
        self.setup_ms, self.symbolic_ms = info[6] / 1e3, info[7] / 1e3

    def panel_ranges(self, budget_bytes: int):
        bounds = np.zeros(self.clusters + 2, np.int64)
        n = c_i64()
        L.call("spops_lespgemm_plan_panels", self._h, int(budget_bytes), ctypes.byref(n), bounds.ctypes.data)
        return [(int(bounds[i]), int(bounds[i + 1])) for i in range(n.value)]

    def numeric(self, c0: int, c1: int, C: ClusterMatrix | None = None) -> ClusterMatrix:
        """Clusters [c0, c1) of the product; C's buffers are reused when given."""
        h = ctypes.c_void_p(C._h.value if C is not None else None)
        stats = (c_i64 * 16)()
        L.call("spops_lespgemm_plan_numeric", self._h, int(c0), int(c1), ctypes.byref(h), stats, L.current_stream())
        if C is None:
            C = ClusterMatrix(h, self.value_type)
        else:
            C._refresh()
        C.stats = _stats(stats)
        return C

    def close(self):
        if self._h:
            L.lib.spops_lespgemm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cluster_flops(A: Csr, B_rowptr, perm, offsets) -> np.ndarray:
    """Flop of each cluster of C = A * B in get_spgemm_flop's sense (src/spgemm.cpp:24-42:
    2 * the B-row lengths of the cluster rows' entries), from host arrays."""
    blen = np.diff(np.asarray(B_rowptr, dtype=np.int64))
    rp = np.asarray(A.rowptr, dtype=np.int64)
    col = np.asarray(A.col, dtype=np.int64)
    rowf = np.zeros(A.n_rows, np.int64)
    nzr = np.diff(rp) > 0
    if A.nnz:
        inb = col < len(blen)
        w = np.where(inb, blen[np.minimum(col, max(len(blen) - 1, 0))], 0)
        rowf[nzr] = np.add.reduceat(w, rp[:-1][nzr])
    perm = np.asarray(perm, dtype=np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)
    csz = np.diff(offsets)
    out = np.zeros(len(csz), np.int64)
    ne = csz > 0
    out[ne] = 2 * np.add.reduceat(rowf[perm], offsets[:-1][ne])
    return out


def sample_ranges(cf, share: float, starts=(0.0, 0.5)):
    """Cluster ranges [c0, c1) of about `share` of the total flop each, starting at the given
    fractions of the flop prefix (the sampled panels timed against the reference)."""
    cum = np.cumsum(np.asarray(cf, dtype=np.int64))
    tot = int(cum[-1]) if len(cum) else 0
    out = []
    for start in starts:
        c0 = int(np.searchsorted(cum, start * tot, side="right"))
        base = cum[c0 - 1] if c0 else 0
        c1 = int(np.searchsorted(cum, base + share * tot)) + 1
        out.append((c0, min(c1, len(cum))))
    return out


def _take_csr(n_rows, n_cols, nnz, drp, dcol, dval, vt) -> Csr:
    """Host copy of three device arrays the library allocated, which are then freed."""
    try:
        rp = np.zeros(n_rows + 1, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz, vt)
        L.call("spops_memcpy_dtoh", rp.ctypes.data, drp, rp.nbytes, None)
        if nnz:
            L.call("spops_memcpy_dtoh", col.ctypes.data, dcol, col.nbytes, None)
            L.call("spops_memcpy_dtoh", val.ctypes.data, dval, val.nbytes, None)
    finally:
        for p in (drp, dcol, dval):
            L.lib.spops_free(p)
    return Csr(n_rows, n_cols, rp, col, val)


def spgemm_rows(dA: DeviceCsr, dB: DeviceCsr, stats=None) -> Csr:
    """Row-wise C = A * B (the reference's LeSpGEMM, src/spgemm.cpp:684-705): columns ascending
    in every row, zero sums kept.  A and B must share the value type."""
    if dA.n_cols != dB.n_rows:
        raise ValueError("Error: A.cols != B.rows")
    if dA.value_type != dB.value_type:
        raise ValueError("A and B must have the same value type")
    nnz = c_i64()
    drp, dcol, dval = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    st = (c_i64 * 16)()
    L.call("spops_spgemm_csr", dA.n_rows, dA.n_cols, dA.nnz, dA.rowptr.data_ptr(), dA.col.data_ptr(),
           dA.val.data_ptr(), dB.n_cols, dB.nnz, dB.rowptr.data_ptr(), dB.col.data_ptr(), dB.val.data_ptr(),
           dA.value_type, ctypes.byref(nnz), ctypes.byref(drp), ctypes.byref(dcol), ctypes.byref(dval), st,
           L.current_stream())
    if stats is not None:
        stats.update(_stats(st))
    vt = np.float64 if dA.value_type == L.VALUE_F64 else np.float32
    return _take_csr(dA.n_rows, dB.n_cols, nnz.value, drp, dcol, dval, vt)


def topk_jaccard(dA: DeviceCsr, top_k: int = 7, precision: int = 64, budget_bytes: int = 0) -> Csr:
    """HashSpGEMMTopK (src/spgemm.cpp:707-842): row i keeps the top_k rows j != i sharing a
    column with it, by Jaccard score descending then j ascending; values are the scores
    (float64, or float32 with precision 32)."""
    nnz = c_i64()
    drp, dcol, dval = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    vt = L.VALUE_F64 if precision == 64 else L.VALUE_F32
    L.call("spops_spgemm_topk_jaccard", dA.n_rows, dA.n_cols, dA.nnz, dA.rowptr.data_ptr(), dA.col.data_ptr(), vt,
           top_k, budget_bytes, ctypes.byref(nnz), ctypes.byref(drp), ctypes.byref(dcol), ctypes.byref(dval),
           L.current_stream())
    return _take_csr(dA.n_rows, dA.n_rows, nnz.value, drp, dcol, dval, np.float64 if precision == 64 else np.float32)


def spgemm_flop(dA: DeviceCsr, dB: DeviceCsr) -> int:
    f = c_i64()
    L.call("spops_spgemm_flop", dA.nnz, dA.col.data_ptr(), dB.n_rows, dB.rowptr.data_ptr(), ctypes.byref(f),
           L.current_stream())
    return f.value


def lespgemm(A: Csr, B: Csr, k: int = 64, bands: int = 16, seed: int = 7, cluster_size: int = 8, hc_v: int = 0,
             device="cuda:0", precision: int = 64, iterations: int = 0):
    """Full test_spgemm_hc_lsh pipeline; returns host arrays and phase timings (ms).

    hc_v: 0 = constrained LSH + v0_fast, 1 = full LSH + v1, 2 = full LSH + v0_fast,
          3 = constrained LSH + the approximate GPU clustering (cluster_rows_gpu; no reference counterpart).
    iterations > 0 additionally times that many LeSpGEMM runs (symbolic + malloc + numeric).
    """
    if A.n_cols != B.n_rows:
        raise ValueError("Error: A.cols != B.rows")
    if k <= 0 or bands <= 0 or k % bands != 0:
        raise ValueError(f"Error: k must be positive and k % bands == 0. k={k} bands={bands}")
    vdt = torch.float64 if precision == 64 else torch.float32
    dA = DeviceCsr(A, device, vdt)
    dB = dA if B is A else DeviceCsr(B, device, vdt)
    t = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if hc_v == 3:
        pi = pj = ps = None
        perm, offsets, rounds, _ = cluster_rows_gpu(dA, k, bands, seed, True, cluster_size,
                                                    score_f32=(precision == 32))
        t["genpairs_hc_ms"] = (time.perf_counter() - t0) * 1e3
        t["hc_rounds"] = rounds
    else:
        pi, pj, ps = lsh_candidate_pairs(dA, k, bands, seed, constrain=(hc_v == 0), score_f32=(precision == 32))
        t["genpairs_ms"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        perm, offsets = hierarchical_clustering(A, pi, pj, ps, cluster_size, version=1 if hc_v == 1 else 0)
        t["hc_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    Acl = ClusterMatrix.build(dA, perm, offsets)
    torch.cuda.synchronize()
    t["conversion_ms"] = (time.perf_counter() - t0) * 1e3
    Ccl = Acl.multiply(dB)
    runs = []
    for _ in range(iterations):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        C2 = Acl.multiply(dB)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t0) * 1e3)
        C2.close()
    C = Ccl.to_csr(Acl)
    nz, mn, mx = Ccl.symbolic()
    out = dict(pi=pi, pj=pj, ps=ps, perm=perm, offsets=offsets, nz=nz, cmin=mn, cmax=mx,
               C_rowptr=C.rowptr, C_col=C.col, C_val=C.val, flop=spgemm_flop(dA, dB), stats=Ccl.stats,
               nclusters=len(offsets) - 1, nnzc_A=Acl.nnzc, nnzv_A=Acl.nnzv, times=t, runs_ms=runs)
    Ccl.close()
    Acl.close()
    return out
