// Device cluster matrix (CSR_VlengthCluster) and SpGEMM constants.
#pragma once

#include "common.hpp"

namespace spops {

constexpr int kSmallFlop = 1024;                 // products of a cluster handled in LDS
constexpr int kSmallTable = 2048;                // LDS hash slots (load factor <= 1/2)
constexpr int kSmallAcc = 4096;                  // LDS accumulator entries (nz * csz)
constexpr int kLargeTileBits = 1 << 18;         // columns per LDS bitmap pass (32 KiB)
constexpr int kLargeBlock = 512;                 // threads of a large-cluster numeric workgroup
constexpr int kLargeAccEntries = 13824;          // LDS accumulator entries of a numeric chunk
constexpr int kLargeHash = 4096;                 // LDS hash slots of a chunk's columns
constexpr int kLargeCursors = 2048;              // A columns whose B-row position is kept in LDS
constexpr int kLargeGroupLanes = 8;              // lanes scanning one B row in the numeric pass
constexpr int kMaxClusterRows = 16;              // A values of a column held in registers per pass
constexpr int kSpgemmStats = 16;

}  // namespace spops

struct spops_cluster_s {
  int64_t csr_rows = 0, rows = 0, cols = 0, nnzc = 0, nnzv = 0, max_cluster_sz = 0;
  int value_type = SPOPS_VALUE_F64;
  spops::DevBuf cluster_sz;          // int32 [rows]
  spops::DevBuf rowptr, rowptr_val;  // int64 [rows + 1]
  spops::DevBuf colids;              // int32 [nnzc]
  spops::DevBuf values;              // V [nnzv]
  // row clustering (A): perm (int64 [csr_rows]), offsets (int64 [rows+1]), and per
  // reordered row its cluster / local row (int32)
  spops::DevBuf perm, offsets, cid, lrow;
  // per-cluster symbolic results (C): nz, min and max column (int64 [rows])
  spops::DevBuf nzc, cmin, cmax;
};

namespace spops {
void cluster_build(spops_cluster_s &A, int64_t n, int64_t n_cols, int64_t nnz, const int64_t *rowptr, const int64_t *col,
                   const void *values, int vt, const int64_t *h_perm, int64_t nc, const int64_t *h_offsets, hipStream_t s);
void lespgemm(const spops_cluster_s &A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *brp64,
              const int64_t *bcol64, const void *bval, spops_cluster_s &C, int64_t stats[kSpgemmStats], hipStream_t s);
void cluster_to_csr(const spops_cluster_s &C, const spops_cluster_s &A, int64_t &nnz, DevBuf &rowptr, DevBuf &col,
                    DevBuf &val, hipStream_t s);
}  // namespace spops
