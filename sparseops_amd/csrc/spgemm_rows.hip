// Row-wise SpGEMM and the A * A^T top-k Jaccard candidate front end (device).
//
//   spops_spgemm_csr             C = A * B in CSR: the reference's row-wise LeSpGEMM
//                                (src/spgemm.cpp:684-705; kernel flags 1 hash, 2 array, 3 SPA,
//                                :45-236 -- all three give the same C)
//   spops_spgemm_topk_jaccard    HashSpGEMMTopK (src/spgemm.cpp:707-842): per row i of A the
//                                rows j != i sharing a column, scored by the Jaccard index
//                                |A_i & A_j| / |A_i | A_j| of the column patterns, the top_k by
//                                (score descending, column ascending) (src/spgemm_hash.cpp:
//                                99-108, 169-230)
//
// A row is a cluster of one row: the product runs on the cluster pipeline (bitmap symbolic
// pass, LDS-hash / chunked / dense numeric kernels chosen per row by its work), whose output
// for singleton clusters is C's CSR with ascending columns.  A's duplicate entries are merged
// first (a radix sort of (row, column) keys, values summed in entry order): the row-wise
// kernels count every stored entry, the cluster format holds one value per (row, column).  The top-k front end builds the
// pattern transpose A^T by a radix sort of (column, row) keys, computes the intersection
// counts |A_i & A_j| as the product of the 0/1 patterns in panels of rows (bounded memory),
// and selects each row's top_k with one wave per row: k rounds, each taking the best entry
// ranked after the previous round's pick, so a row needs no sort and no scratch.
#include <algorithm>
#include <memory>
#include <vector>

#include "primitives.hpp"
#include "spgemm_internal.hpp"

namespace spops {
namespace {

__global__ __launch_bounds__(kBlock) void k_widen_col(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) out[i] = in[i];
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_fill_one(V *__restrict__ v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) v[i] = V(1);
}

// (column << 32 | row) of every entry, one wave per row; the column count of every column
__global__ __launch_bounds__(kBlock) void k_transpose_keys(const int64_t *__restrict__ rowptr, const int64_t *__restrict__ col,
                                                           int64_t n_rows, uint64_t *__restrict__ keys,
                                                           unsigned long long *__restrict__ ccount) {
  const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t r = blockIdx.x * (int64_t)(kBlock / kWave) + wave_id(); r < n_rows; r += waves)
    for (int64_t e = rowptr[r] + lane_id(); e < rowptr[r + 1]; e += kWave) {
      const int64_t c = col[e];
      keys[e] = (uint64_t)c << 32 | (uint64_t)r;
      atomicAdd(&ccount[c], 1ull);
    }
}

__global__ __launch_bounds__(kBlock) void k_key_low(const uint64_t *__restrict__ keys, int64_t n, int64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    out[i] = (int64_t)(keys[i] & 0xFFFFFFFFull);
}

// (score, column) of a candidate ranked before (s2, c2): higher score, then lower column
template <typename V>
__device__ __forceinline__ bool ranks_before(V s1, int c1, V s2, int c2) {
  return s1 > s2 || (s1 == s2 && c1 < c2);
}

// Top-k of rows [r0, r0 + nr) from their intersection counts (the product of the patterns,
// rows local to the panel: crp[0..nr]).  One wave per row; round t picks the entry ranked
// first among those ranked after round t-1's pick.  out slots: row r's picks at r * k.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_topk_rows(int64_t r0, int64_t nr, const int64_t *__restrict__ crp,
                                                      const int32_t *__restrict__ ccol, const V *__restrict__ cval,
                                                      const int64_t *__restrict__ arp, int64_t k,
                                                      int32_t *__restrict__ ocol, V *__restrict__ oval,
                                                      int64_t *__restrict__ ocnt) {
  const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
  const int lane = lane_id();
  for (int64_t lr = blockIdx.x * (int64_t)(kBlock / kWave) + wave_id(); lr < nr; lr += waves) {
    const int64_t r = r0 + lr;
    const int64_t beg = crp[lr], end = crp[lr + 1];
    const int64_t li = arp[r + 1] - arp[r];
    V prev_s = V(0);
    int prev_c = -1;
    int64_t t = 0;
    for (; t < k; ++t) {
      V best_s = V(-1);                        // below every score (scores are >= 0)
      int best_c = INT32_MAX;
      for (int64_t e = beg + lane; e < end; e += kWave) {
        const int c = ccol[e];
        if (c == (int)r) continue;             // the row itself
        const V inter = cval[e];
        const int64_t u = li + (arp[c + 1] - arp[c]) - (int64_t)inter;
        const V s = u == 0 ? V(0) : V(1) * inter / (V)u;
        if (t > 0 && !ranks_before(prev_s, prev_c, s, c)) continue;   // taken by an earlier round
        if (ranks_before(s, c, best_s, best_c)) { best_s = s; best_c = c; }
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        const V os = __shfl_xor(best_s, off, kWave);
        const int oc = __shfl_xor(best_c, off, kWave);
        if (ranks_before(os, oc, best_s, best_c)) { best_s = os; best_c = oc; }
      }
      if (best_s < V(0)) break;                // no entry left
      if (lane == 0) {
        ocol[r * k + t] = best_c;
        oval[r * k + t] = best_s;
      }
      prev_s = best_s;
      prev_c = best_c;
    }
    if (lane == 0) ocnt[r] = t;
  }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_topk_compact(int64_t n_rows, int64_t k, const int32_t *__restrict__ scol,
                                                         const V *__restrict__ sval, const int64_t *__restrict__ rowptr,
                                                         int64_t *__restrict__ col, V *__restrict__ val) {
  const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t r = blockIdx.x * (int64_t)(kBlock / kWave) + wave_id(); r < n_rows; r += waves)
    for (int64_t t = lane_id(); t < rowptr[r + 1] - rowptr[r]; t += kWave) {
      col[rowptr[r] + t] = scol[r * k + t];
      val[rowptr[r] + t] = sval[r * k + t];
    }
}

// (row << 32 | column) and the entry index of every entry, one wave per row
__global__ __launch_bounds__(kBlock) void k_entry_keys(const int64_t *__restrict__ rowptr, const int64_t *__restrict__ col,
                                                       int64_t n_rows, uint64_t *__restrict__ keys,
                                                       int64_t *__restrict__ idx) {
  const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t r = blockIdx.x * (int64_t)(kBlock / kWave) + wave_id(); r < n_rows; r += waves)
    for (int64_t e = rowptr[r] + lane_id(); e < rowptr[r + 1]; e += kWave) {
      keys[e] = (uint64_t)r << 32 | (uint64_t)col[e];
      idx[e] = e;
    }
}

// 1 where a sorted key starts a run of equal keys; the row of each run head counted
__global__ __launch_bounds__(kBlock) void k_run_heads(const uint64_t *__restrict__ keys, int64_t n, int64_t *__restrict__ head,
                                                      unsigned long long *__restrict__ rowcnt) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const bool h = i == 0 || keys[i] != keys[i - 1];
    head[i] = h ? 1 : 0;
    if (h) atomicAdd(&rowcnt[keys[i] >> 32], 1ull);
  }
}

// the run starting at each head: its column, and its values summed in entry order
template <typename V>
__global__ __launch_bounds__(kBlock) void k_run_sums(const uint64_t *__restrict__ keys, const int64_t *__restrict__ idx,
                                                     const int64_t *__restrict__ pos, int64_t n,
                                                     const V *__restrict__ val, int64_t *__restrict__ ucol,
                                                     V *__restrict__ uval) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    if (!(i == 0 || keys[i] != keys[i - 1])) continue;
    V sum = val[idx[i]];
    for (int64_t f = i + 1; f < n && keys[f] == keys[i]; ++f) sum += val[idx[f]];
    ucol[pos[i]] = (int64_t)(keys[i] & 0xFFFFFFFFull);
    uval[pos[i]] = sum;
  }
}

unsigned grid_for(int64_t items, int per_block) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + per_block - 1) / per_block, 65536));
}

// A with the duplicate entries of each row merged (values summed in entry order) and columns
// ascending: the row-wise product counts every stored entry, while the cluster format keeps one
// value per (row, column)
template <typename V>
void merge_duplicates(int64_t n_rows, int64_t nnz, const int64_t *rowptr, const int64_t *col, const V *val,
                      DevBuf &urp, DevBuf &ucol, DevBuf &uval, int64_t &unnz, hipStream_t s) {
  DevBuf keys, idx, head, pos, rowcnt;
  keys.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  idx.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  head.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  pos.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  rowcnt.alloc((size_t)(n_rows + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(rowcnt.p, 0, (size_t)(n_rows + 1) * 8, s));
  hipLaunchKernelGGL(k_entry_keys, dim3(grid_for(n_rows, kBlock / kWave)), dim3(kBlock), 0, s, rowptr, col, n_rows,
                     keys.as<uint64_t>(), idx.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  radix_sort_u64(keys.as<uint64_t>(), idx.p, 8, nnz, 0, 32 + std::max(1, bit_width_u64((uint64_t)n_rows)), s);
  hipLaunchKernelGGL(k_run_heads, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(), nnz,
                     head.as<int64_t>(), rowcnt.as<unsigned long long>());
  SPOPS_HIP_CHECK(hipGetLastError());
  exclusive_scan_i64(head.as<int64_t>(), pos.as<int64_t>(), nnz, s, nullptr);
  urp.alloc((size_t)(n_rows + 1) * 8);
  exclusive_scan_i64(rowcnt.as<int64_t>(), urp.as<int64_t>(), n_rows + 1, s, &unnz);
  ucol.alloc((size_t)std::max<int64_t>(unnz, 1) * 8);
  uval.alloc((size_t)std::max<int64_t>(unnz, 1) * sizeof(V));
  hipLaunchKernelGGL(k_run_sums<V>, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(),
                     idx.as<int64_t>(), pos.as<int64_t>(), nnz, val, ucol.as<int64_t>(), uval.as<V>());
  SPOPS_HIP_CHECK(hipGetLastError());
}

void merge_duplicates_vt(int vt, int64_t n_rows, int64_t nnz, const int64_t *rowptr, const int64_t *col,
                         const void *val, DevBuf &urp, DevBuf &ucol, DevBuf &uval, int64_t &unnz, hipStream_t s) {
  if (vt == SPOPS_VALUE_F64)
    merge_duplicates<double>(n_rows, nnz, rowptr, col, static_cast<const double *>(val), urp, ucol, uval, unnz, s);
  else
    merge_duplicates<float>(n_rows, nnz, rowptr, col, static_cast<const float *>(val), urp, ucol, uval, unnz, s);
}

// Singleton-cluster matrix of a CSR matrix on the device (cluster c = row c)
void rows_as_clusters(spops_cluster_s &A, int64_t n, int64_t n_cols, int64_t nnz, const int64_t *rowptr,
                      const int64_t *col, const void *values, int vt, hipStream_t s) {
  std::vector<int64_t> iota((size_t)n + 1);
  for (int64_t i = 0; i <= n; ++i) iota[(size_t)i] = i;
  cluster_build(A, n, n_cols, nnz, rowptr, col, values, vt, iota.data(), n, iota.data(), s);
}

// A^T of a 0/1 pattern: rowptr (n_cols + 1) and rows of each column ascending
void pattern_transpose(int64_t n_rows, int64_t n_cols, int64_t nnz, const int64_t *rowptr, const int64_t *col,
                       DevBuf &trp, DevBuf &tcol, hipStream_t s) {
  trp.alloc((size_t)(n_cols + 1) * 8);
  tcol.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  DevBuf keys, cnt;
  keys.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  cnt.alloc((size_t)(n_cols + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(cnt.p, 0, (size_t)(n_cols + 1) * 8, s));
  if (n_rows > 0 && nnz > 0)
    hipLaunchKernelGGL(k_transpose_keys, dim3(grid_for(n_rows, kBlock / kWave)), dim3(kBlock), 0, s, rowptr, col, n_rows,
                       keys.as<uint64_t>(), cnt.as<unsigned long long>());
  SPOPS_HIP_CHECK(hipGetLastError());
  exclusive_scan_i64(cnt.as<int64_t>(), trp.as<int64_t>(), n_cols + 1, s, nullptr);
  if (nnz > 0) {
    radix_sort_u64(keys.as<uint64_t>(), nullptr, 8, nnz, 0, 32 + std::max(1, bit_width_u64((uint64_t)n_cols)), s);
    hipLaunchKernelGGL(k_key_low, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(), nnz,
                       tcol.as<int64_t>());
  }
  SPOPS_HIP_CHECK(hipGetLastError());
}

template <typename V>
void topk_jaccard(int64_t n_rows, int64_t n_cols, int64_t nnz, const int64_t *rowptr, const int64_t *col, int vt,
                  int64_t k, int64_t budget, int64_t &nnz_out, DevBuf &orp, DevBuf &ocol, DevBuf &oval, hipStream_t s) {
  const size_t vb = sizeof(V);
  // k picks per row at most, and never more than the other rows
  const int64_t kk = std::max<int64_t>(0, std::min<int64_t>(k, std::max<int64_t>(n_rows - 1, 0)));
  if (nnz == 0 || kk == 0) {                 // no candidates: every row empty
    orp.alloc((size_t)(n_rows + 1) * 8);
    SPOPS_HIP_CHECK(hipMemsetAsync(orp.p, 0, (size_t)(n_rows + 1) * 8, s));
    ocol.alloc(8);
    oval.alloc(8);
    nnz_out = 0;
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  DevBuf ones;
  ones.alloc((size_t)std::max<int64_t>(nnz, 1) * vb);
  if (nnz > 0) hipLaunchKernelGGL(k_fill_one<V>, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, ones.as<V>(), nnz);
  SPOPS_HIP_CHECK(hipGetLastError());
  spops_cluster_s Acl;
  {
    DevBuf urp, ucol, uval;          // A's pattern with duplicates merged: values = multiplicities
    int64_t unnz = 0;
    merge_duplicates<V>(n_rows, nnz, rowptr, col, ones.as<V>(), urp, ucol, uval, unnz, s);
    rows_as_clusters(Acl, n_rows, n_cols, unnz, urp.as<int64_t>(), ucol.as<int64_t>(), uval.p, vt, s);
  }
  DevBuf trp, tcol;
  pattern_transpose(n_rows, n_cols, nnz, rowptr, col, trp, tcol, s);
  spops_spgemm_plan_s P;
  spgemm_plan_create(Acl, n_cols, n_rows, nnz, trp.as<int64_t>(), tcol.as<int64_t>(), ones.p, P, s);
  SPOPS_REQUIRE((double)n_rows * (double)kk * (4.0 + vb) <= (double)((int64_t)16 << 30), SPOPS_ERROR_LIMIT,
                "top_k slots (n_rows * top_k = %lld) exceed 16 GiB", (long long)(n_rows * kk));
  DevBuf scol, sval, cnt;
  scol.alloc((size_t)std::max<int64_t>(n_rows * kk, 1) * 4);
  sval.alloc((size_t)std::max<int64_t>(n_rows * kk, 1) * vb);
  cnt.alloc((size_t)(n_rows + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(cnt.p, 0, (size_t)(n_rows + 1) * 8, s));
  {
    // panels of rows whose intersection counts (column ids + values) fit the budget
    if (budget <= 0) {
      size_t free_b = 0, total_b = 0;
      SPOPS_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
      budget = std::max<int64_t>((int64_t)free_b / 4, (int64_t)64 << 20);
    }
    spops_cluster_s Cp;
    int64_t st[kSpgemmStats];
    int64_t c0 = 0;
    while (c0 < n_rows) {
      int64_t c1 = c0, bytes = 0;
      while (c1 < n_rows && (c1 == c0 || bytes + P.hnz[(size_t)c1] * (int64_t)(4 + vb) <= budget))
        bytes += P.hnz[(size_t)c1++] * (int64_t)(4 + vb);
      spgemm_plan_numeric(P, c0, c1, Cp, st, s);
      hipLaunchKernelGGL(k_topk_rows<V>, dim3(grid_for(c1 - c0, kBlock / kWave)), dim3(kBlock), 0, s, c0, c1 - c0,
                         Cp.rowptr.as<int64_t>(), Cp.colids.as<int32_t>(), Cp.values.as<V>(), rowptr, kk,
                         scol.as<int32_t>(), sval.as<V>(), cnt.as<int64_t>());
      SPOPS_HIP_CHECK(hipGetLastError());
      c0 = c1;
    }
  }
  orp.alloc((size_t)(n_rows + 1) * 8);
  exclusive_scan_i64(cnt.as<int64_t>(), orp.as<int64_t>(), n_rows + 1, s, &nnz_out);
  ocol.alloc((size_t)std::max<int64_t>(nnz_out, 1) * 8);
  oval.alloc((size_t)std::max<int64_t>(nnz_out, 1) * vb);
  if (nnz_out > 0)
    hipLaunchKernelGGL(k_topk_compact<V>, dim3(grid_for(n_rows, kBlock / kWave)), dim3(kBlock), 0, s, n_rows, kk,
                       scol.as<int32_t>(), sval.as<V>(), orp.as<int64_t>(), ocol.as<int64_t>(), oval.as<V>());
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace spops

using namespace spops;

extern "C" int spops_spgemm_csr(int64_t a_rows, int64_t a_cols, int64_t a_nnz, const int64_t *d_arowptr,
                                const int64_t *d_acol, const void *d_aval, int64_t b_cols, int64_t b_nnz,
                                const int64_t *d_browptr, const int64_t *d_bcol, const void *d_bval, int value_type,
                                int64_t *c_nnz, int64_t **d_crowptr, int64_t **d_ccol, void **d_cval, int64_t stats[16],
                                void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(c_nnz && d_crowptr && d_ccol && d_cval, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(value_type == SPOPS_VALUE_F64 || value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "values must be F64 or F32");
    SPOPS_REQUIRE(a_rows >= 0 && a_cols >= 0 && a_nnz >= 0 && b_cols >= 0 && b_nnz >= 0, SPOPS_ERROR_INVALID_VALUE,
                  "negative size");
    SPOPS_REQUIRE(a_rows < ((int64_t)1 << 31) && a_cols < ((int64_t)1 << 31) && b_cols < ((int64_t)1 << 31),
                  SPOPS_ERROR_LIMIT, "rows and columns must be < 2^31");
    SPOPS_REQUIRE((a_rows == 0 || (d_arowptr && (a_nnz == 0 || (d_acol && d_aval)))) &&
                      (a_cols == 0 || (d_browptr && (b_nnz == 0 || (d_bcol && d_bval)))),
                  SPOPS_ERROR_INVALID_VALUE, "null input array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf rp, col, val;
    int64_t st[kSpgemmStats] = {0};
    if (a_rows == 0 || a_nnz == 0 || b_nnz == 0) {     // no products: every row of C empty
      rp.alloc((size_t)(a_rows + 1) * 8);
      SPOPS_HIP_CHECK(hipMemsetAsync(rp.p, 0, (size_t)(a_rows + 1) * 8, s));
      SPOPS_HIP_CHECK(hipStreamSynchronize(s));
      col.alloc(8);
      val.alloc(8);
      *c_nnz = 0;
    } else {
      spops_cluster_s Acl, C;
      {
        DevBuf urp, ucol, uval;
        int64_t unnz = 0;
        merge_duplicates_vt(value_type, a_rows, a_nnz, d_arowptr, d_acol, d_aval, urp, ucol, uval, unnz, s);
        rows_as_clusters(Acl, a_rows, a_cols, unnz, urp.as<int64_t>(), ucol.as<int64_t>(), uval.p, value_type, s);
      }
      lespgemm(Acl, a_cols, b_cols, b_nnz, d_browptr, d_bcol, d_bval, C, st, s);
      *c_nnz = C.nnzc;
      rp = std::move(C.rowptr);
      val = std::move(C.values);
      col.alloc((size_t)std::max<int64_t>(C.nnzc, 1) * 8);
      if (C.nnzc > 0)
        hipLaunchKernelGGL(k_widen_col, dim3(grid_for(C.nnzc, kBlock)), dim3(kBlock), 0, s, C.colids.as<int32_t>(),
                           C.nnzc, col.as<int64_t>());
      SPOPS_HIP_CHECK(hipGetLastError());
      SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (stats)
      for (int i = 0; i < 16; ++i) stats[i] = i < kSpgemmStats ? st[i] : 0;
    *d_crowptr = static_cast<int64_t *>(rp.p); rp.p = nullptr;
    *d_ccol = static_cast<int64_t *>(col.p); col.p = nullptr;
    *d_cval = val.p; val.p = nullptr;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spgemm_topk_jaccard(int64_t n_rows, int64_t n_cols, int64_t nnz, const int64_t *d_rowptr,
                                         const int64_t *d_col, int value_type, int64_t top_k, int64_t budget_bytes,
                                         int64_t *c_nnz, int64_t **d_crowptr, int64_t **d_ccol, void **d_cval,
                                         void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(c_nnz && d_crowptr && d_ccol && d_cval, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(value_type == SPOPS_VALUE_F64 || value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "values must be F64 or F32");
    SPOPS_REQUIRE(n_rows >= 0 && n_cols >= 0 && nnz >= 0 && top_k >= 0, SPOPS_ERROR_INVALID_VALUE,
                  "negative size or top_k");
    SPOPS_REQUIRE(n_rows < ((int64_t)1 << 31) && n_cols < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT,
                  "rows and columns must be < 2^31");
    SPOPS_REQUIRE(n_rows == 0 || (d_rowptr && (nnz == 0 || d_col)), SPOPS_ERROR_INVALID_VALUE, "null input array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf rp, col, val;
    if (value_type == SPOPS_VALUE_F64)
      topk_jaccard<double>(n_rows, n_cols, nnz, d_rowptr, d_col, value_type, top_k, budget_bytes, *c_nnz, rp, col, val,
                           s);
    else
      topk_jaccard<float>(n_rows, n_cols, nnz, d_rowptr, d_col, value_type, top_k, budget_bytes, *c_nnz, rp, col, val,
                          s);
    *d_crowptr = static_cast<int64_t *>(rp.p); rp.p = nullptr;
    *d_ccol = static_cast<int64_t *>(col.p); col.p = nullptr;
    *d_cval = val.p; val.p = nullptr;
    return SPOPS_SUCCESS;
  });
}
