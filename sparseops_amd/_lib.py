"""ctypes binding of libsparseops_hip.so (the C ABI in include/spops_hip.h).

torch is imported first so that the library binds to the same HIP runtime that torch
loaded (one runtime per process); tensors are passed by device pointer.  There is no
fallback: if the shared library is missing, importing raises.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library then shares)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsparseops_hip.so")

SUCCESS = 0
NOT_SUPPORTED = 4
INDEX_I32, INDEX_I64 = 0, 1
VALUE_F64, VALUE_F32, VALUE_BF16 = 0, 1, 2
FORMAT_CSR, FORMAT_SELL, FORMAT_ELL, FORMAT_BSR = 0, 1, 2, 3
(TUNE_SLICE_BYTES, TUNE_DENSE_PRODUCTS, TUNE_SPMV_LAYOUT, TUNE_SWEEP_WORKGROUPS, TUNE_SWEEP_MAX_ROWS,
 TUNE_BSR_MFMA, TUNE_TASK_PRODUCTS, TUNE_WIDE_CURSORS, TUNE_HEAVY_PRODUCTS, TUNE_SWEEP_SPLIT) = range(10)

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_dbl = ctypes.c_double

# (name, restype, argtypes) for every symbol of include/spops_hip.h and the other headers.
_SIGNATURES = [
    ("spops_status_string", ctypes.c_char_p, [c_int]),
    ("spops_last_error", ctypes.c_char_p, []),
    ("spops_version", c_int, [c_vp, c_vp, c_vp]),
    ("spops_device_count", c_int, [c_vp]),
    ("spops_set_device", c_int, [c_int]),
    ("spops_malloc", c_int, [c_vp, ctypes.c_size_t]),
    ("spops_free", c_int, [c_vp]),
    ("spops_memcpy_htod", c_int, [c_vp, c_vp, ctypes.c_size_t, c_vp]),
    ("spops_memcpy_dtoh", c_int, [c_vp, c_vp, ctypes.c_size_t, c_vp]),
    ("spops_memcpy_dtod", c_int, [c_vp, c_vp, ctypes.c_size_t, c_vp]),
    ("spops_memset", c_int, [c_vp, c_int, ctypes.c_size_t, c_vp]),
    ("spops_stream_synchronize", c_int, [c_vp]),
    ("spops_stream_create", c_int, [c_vp]),
    ("spops_stream_destroy", c_int, [c_vp]),
    ("spops_event_create", c_int, [c_vp]),
    ("spops_event_destroy", c_int, [c_vp]),
    ("spops_event_record", c_int, [c_vp, c_vp]),
    ("spops_event_elapsed_ms", c_int, [c_vp, c_vp, c_vp]),
    ("spops_csr_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int]),
    ("spops_csr_create_layout", c_int, [c_vp, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int, c_int]),
    ("spops_sell_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int, c_i64, c_i64]),
    ("spops_sell_create_from_chunks", c_int,
     [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_int, c_int]),
    ("spops_ell_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int]),
    ("spops_bsr_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int, c_i64, c_i64]),
    ("spops_bsr_create_from_blocks", c_int,
     [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_int]),
    ("spops_spmv", c_int, [c_vp, c_dbl, c_vp, c_dbl, c_vp, c_vp]),
    ("spops_spmm", c_int, [c_vp, c_i64, c_dbl, c_vp, c_dbl, c_vp, c_vp]),
    ("spops_spmm_typed", c_int, [c_vp, c_i64, c_dbl, c_vp, c_int, c_dbl, c_vp, c_int, c_vp]),
    ("spops_spmm_values", c_int, [c_vp, c_vp, c_i64, c_dbl, c_vp, c_int, c_dbl, c_vp, c_int, c_vp]),
    ("spops_spmm_typed_noalloc", c_int, [c_vp, c_i64, c_dbl, c_vp, c_int, c_dbl, c_vp, c_int, c_vp]),
    ("spops_sddmm", c_int, [c_vp, c_i64, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_vp]),
    ("spops_spmm_sddmm_transposed", c_int, [c_vp, c_i64, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp]),
    ("spops_csr_transpose", c_int, [c_vp, c_vp]),
    ("spops_csr_set_values", c_int, [c_vp, c_vp, c_int, c_vp]),
    ("spops_matrix_features", c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_vp]),
    ("spops_batch_create", c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_int]),
    ("spops_batch_info", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_batch_plan", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_batch_spmm", c_int, [c_vp, ctypes.c_float, c_vp, ctypes.c_float, c_vp, c_vp]),
    ("spops_batch_capture", c_int, [c_vp, ctypes.c_float, c_vp, ctypes.c_float, c_vp, c_vp]),
    ("spops_batch_replay", c_int, [c_vp, c_vp]),
    ("spops_batch_destroy", c_int, [c_vp]),
    ("spops_matrix_info", c_int, [c_vp, c_vp]),
    ("spops_matrix_layout", c_int, [c_vp, c_int, c_vp]),
    ("spops_set_tuning", c_int, [c_int, c_i64, c_vp]),
    ("spops_matrix_destroy", c_int, [c_vp]),
    ("spops_minhash_signatures", c_int, [c_vp, c_vp, c_int, c_i64, c_int, ctypes.c_uint64, c_vp, c_vp]),
    ("spops_lsh_candidate_pairs", c_int,
     [c_vp, c_vp, c_int, c_i64, c_int, c_int, ctypes.c_uint64, c_int, c_int, c_vp, c_vp]),
    ("spops_lsh_pairs_from_signatures", c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp]),
    ("spops_pairs_count", c_int, [c_vp, c_vp]),
    ("spops_pairs_device", c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("spops_pairs_download", c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("spops_pairs_destroy", c_int, [c_vp]),
    ("spops_hierarchical_clustering", c_int,
     [c_int, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_int, c_i64, c_vp, c_vp, c_vp]),
    ("spops_cluster_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_int, c_vp, c_i64, c_vp, c_vp]),
    ("spops_cluster_create_from_host", c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]),
    ("spops_cluster_info", c_int, [c_vp, c_vp]),
    ("spops_cluster_download", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_cluster_symbolic_download", c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("spops_cluster_destroy", c_int, [c_vp]),
    ("spops_lespgemm_plan_create", c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_lespgemm_plan_info", c_int, [c_vp, c_vp]),
    ("spops_lespgemm_plan_panels", c_int, [c_vp, c_i64, c_vp, c_vp]),
    ("spops_lespgemm_plan_numeric", c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    ("spops_lespgemm_plan_destroy", c_int, [c_vp]),
    ("spops_cluster_checksum", c_int, [c_vp, c_vp, c_vp]),
    ("spops_lespgemm", c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_cluster_to_csr", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_spgemm_flop", c_int, [c_i64, c_vp, c_i64, c_vp, c_vp, c_vp]),
    ("spops_hierarchical_clustering_v0", c_int, [c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_int, c_i64, c_vp, c_vp,
                                                 c_vp, c_vp]),
    ("spops_cluster_matching", c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_int, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp,
                                       c_vp]),
    ("spops_spgemm_csr", c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_int,
                                 c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("spops_spgemm_topk_jaccard", c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_int, c_i64, c_i64, c_vp, c_vp, c_vp,
                                          c_vp, c_vp]),
]


class SpopsError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: status {status} ({_status_name(status)}): {detail}")
        self.status = status


def _status_name(s: int) -> str:
    return {0: "success", 1: "invalid value", 2: "out of memory", 3: "HIP error",
            4: "not supported", 5: "size limit"}.get(s, "unknown")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
            f"or make -C sparseops_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def declared_symbols():
    return [n for n, _, _ in _SIGNATURES]


def check(status: int, where: str) -> None:
    if status != SUCCESS:
        raise SpopsError(status, where, lib.spops_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)


def ptr(a) -> int:
    """Device pointer of a torch tensor or host pointer of a numpy array (None -> 0)."""
    if a is None:
        return 0
    if isinstance(a, torch.Tensor):
        return a.data_ptr()
    return a.ctypes.data


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream
