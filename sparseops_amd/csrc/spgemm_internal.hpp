// Device cluster matrix (CSR_VlengthCluster) and SpGEMM constants.
#pragma once

#include "common.hpp"

namespace spops {

constexpr int kSmallFlop = 1024;                 // products of a cluster handled in LDS
constexpr int kSmallTable = 2048;                // LDS hash slots (load factor <= 1/2)
constexpr int kSmallAcc = 4096;                  // LDS accumulator entries (nz * csz)
constexpr int kLargeTileBits = 1 << 18;         // columns per LDS bitmap pass (32 KiB)
constexpr int kLargeBlock = 512;                 // threads of a large-cluster numeric workgroup
constexpr int kLargeAccEntries = 2304;           // LDS accumulator entries of a numeric chunk (CR = 16)
constexpr int kLargeHash = 1024;                 // LDS hash slots of a chunk's columns
constexpr int kLargeCursors = 1016;              // A columns whose B-row position is kept in LDS
constexpr int kLargeCursorsWide = 4096;          // the same in the variant for wider clusters (two workgroups per CU)
constexpr int64_t kWideMinTasks = 8192;          // per-call choice: fewest tasks for that variant (16 per slot)
constexpr int kLargeGroupLanes = 4;              // fewest lanes scanning one B row in the numeric pass
constexpr int64_t kLargeTaskProducts = (int64_t)1 << 15;    // products per numeric task of a split cluster ...
constexpr int64_t kLargeTaskProductsMin = (int64_t)1 << 13; // ... halved down to this while a call has
constexpr int64_t kLargeTasksMin = 32768;                    // fewer tasks than this
// hash slots of a numeric chunk: clusters of at most 4 rows (CR = 4; mostly single rows) take
// twice as many, so a chunk holds up to 1024 output columns: half the chunks, each of which
// visits every row of the cluster once
constexpr int large_hash_slots(int cr) { return cr <= 4 ? 2 * kLargeHash : kLargeHash; }
// output columns of a numeric chunk of a cluster of csz rows
// LDS accumulator entries: 2048 for clusters of at most 4 rows, so that with 1016 cursors the
// CR = 4 kernel takes 40 KiB of LDS and four workgroups (eight waves per SIMD) fit one CU
constexpr int large_acc_entries(int cr) { return cr <= 4 ? 2048 : kLargeAccEntries; }
constexpr int large_chunk_cols(int csz, int cr) {
  return large_acc_entries(cr) / csz < large_hash_slots(cr) / 2 ? large_acc_entries(cr) / csz
                                                                 : large_hash_slots(cr) / 2;
}
constexpr int kMaxClusterRows = 16;              // A values of a column held in registers per pass
// heavy clusters (at least g_heavy_products products): symbolic bitmaps built in parts of
// about kHeavyPartProducts products by several workgroups and kept for the emit pass
constexpr int64_t kHeavyProducts = (int64_t)1 << 20;        // default of g_heavy_products
constexpr int64_t kHeavyPartProducts = (int64_t)1 << 18;
constexpr int64_t kHeavySliceBytes = (int64_t)1 << 30;      // part bitmaps live during the symbolic pass
constexpr int64_t kHeavyBitmapBytes = (int64_t)1 << 28;     // bitmaps a plan keeps
constexpr int kSpgemmStats = 16;
constexpr int64_t kDenseScratch = (int64_t)1 << 30;    // bytes of dense accumulators live at once
constexpr int kDenseTaskCols = 32;                      // A columns per dense-product workgroup

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

// A LeSpGEMM product C = A * B prepared once and computed in cluster ranges ("panels"): B
// with int32 columns and sorted rows, the per-cluster work, the small / large bins and the
// symbolic results (nz, min / max column) of every cluster.  A and B must outlive the plan.
struct spops_spgemm_plan_s {
  const spops_cluster_s *A = nullptr;
  int64_t b_rows = 0, b_cols = 0, b_nnz = 0;
  const int64_t *brp64 = nullptr;
  const void *bval = nullptr;              // B's values, or the sorted copy below
  spops::DevBuf bcol32, bsorted_val;
  int unsorted = 0;
  std::vector<int64_t> hflop, hnz;         // per cluster: products, distinct columns of C
  std::vector<int32_t> hcsz;
  std::vector<int32_t> small, large;       // clusters with work, by bin (large by decreasing work)
  std::vector<char> dense;                 // per cluster: accumulated in a dense row block in memory
  std::vector<int64_t> harp;               // A's cluster column pointers (host copy)
  std::vector<int64_t> hcmin, hcmax;       // per cluster: column range of C
  spops::DevBuf scratch;                   // dense accumulators (kDenseScratch at most)
  spops::DevBuf dsmall;
  spops::DevBuf gcur;                      // int32 [2][A's nnzc]: B-row cursors of wide clusters
  spops::DevBuf nz, cmin, cmax;            // int64 [rows]
  std::vector<int32_t> heavy_id;           // per cluster: index of its kept bitmap, or -1
  spops::DevBuf heavy_bm;                  // uint32 [heavy clusters][heavy_words]
  spops::DevBuf dlight;                    // int32: large clusters without a kept bitmap
  spops::DevBuf dheavy;                    // int2 {cluster, bitmap}: the heavy clusters
  int64_t heavy_words = 0, nheavy = 0, nlight = 0;
  int64_t nnzc = 0, nnzv = 0, products = 0;
  double ms_setup = 0, ms_symbolic = 0;
};

namespace spops {
extern int64_t g_dense_products;   // products from which a cluster accumulates densely (tuning)
extern int64_t g_task_products;    // products per numeric task of a large cluster, 0: chosen per call (tuning)
extern int g_wide_cursors;         // wide-cluster cursor variant: 0 off, 1 on, 2 per call (tuning)
extern int64_t g_heavy_products;   // products from which a large cluster's bitmap is built in parts (tuning)
void spgemm_plan_create(const spops_cluster_s &A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *brp64,
                        const int64_t *bcol64, const void *bval, spops_spgemm_plan_s &P, hipStream_t s);
// C = rows [c0, c1) of the product as a cluster matrix of c1 - c0 clusters; C's buffers are
// reused when large enough
void spgemm_plan_numeric(spops_spgemm_plan_s &P, int64_t c0, int64_t c1, spops_cluster_s &C,
                         int64_t stats[kSpgemmStats], hipStream_t s);
void cluster_build(spops_cluster_s &A, int64_t n, int64_t n_cols, int64_t nnz, const int64_t *rowptr, const int64_t *col,
                   const void *values, int vt, const int64_t *h_perm, int64_t nc, const int64_t *h_offsets, hipStream_t s);
void lespgemm(const spops_cluster_s &A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *brp64,
              const int64_t *bcol64, const void *bval, spops_cluster_s &C, int64_t stats[kSpgemmStats], hipStream_t s);
void cluster_to_csr(const spops_cluster_s &C, const spops_cluster_s &A, int64_t &nnz, DevBuf &rowptr, DevBuf &col,
                    DevBuf &val, hipStream_t s);
}  // namespace spops
