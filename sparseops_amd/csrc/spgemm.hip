// Cluster-wise SpGEMM (LeSpGEMM, kernel 3 semantics) on MI355X.
//
// Reference path (lsl036/SparseOps):
//   csr_reorder_rows + csr_to_vlength_cluster   include/sparse_conversion.h:1641-1976
//   LeSpGEMM_mixed_acc                          src/spgemm.cpp:545-659
//   spgemm_mixed_symbolic_omp_lb                src/spgemm_Vlength_mixed.cpp:197-261
//   classify_clusters                           src/spgemm_Vlength_mixed.cpp:268-319
//   spgemm_mixed_numeric_omp_lb                 src/spgemm_Vlength_mixed.cpp:326-434
//   vlength_cluster2csr (+ inverse permutation) include/sparse_conversion.h:1778-1862,
//                                               test/test_spgemm_hc_lsh.cpp:296-303
//
// Cluster matrix layout (identical meaning to CSR_VlengthCluster,
// include/spgemm_cluster.h:108-143): cluster c holds csz = cluster_sz[c] consecutive rows
// of the reordered matrix, its distinct columns colids[rowptr[c] .. rowptr[c+1]) and a
// dense csz-wide value block per column at values[rowptr_val[c] + t*csz + l].  Here the
// columns of each cluster are ascending (A: the reference iterates an unordered_map; C:
// the reference's sortOutput=1 order), which changes only FP summation order.
//
// GPU design:
//   * cluster build: keys (cluster << 32 | col) of every entry in reordered order, stable
//     radix sort, head flags + scan give colids and the column offsets; values scatter.
//   * per-cluster work flop_c = sum over its columns of nnz(B row) bins the clusters:
//       small  (flop_c <= kSmallFlop, output block fits LDS): one workgroup, LDS hash
//              table of columns, LDS bitonic sort, LDS f64 accumulator, coalesced store;
//       large  : one workgroup per cluster, all in LDS: bitmap passes over the B rows give
//              the count / min / max and then C's sorted columns; the numeric pass takes
//              those columns in chunks that fit an LDS accumulator (LDS hash column -> rank)
//              and reads only the B entries inside each chunk's column range (B rows are
//              sorted; an unsorted B is sorted into a copy first).
//   * symbolic (counts, min/max column) and numeric (columns + values) are separate
//     passes; C's offsets come from a device scan in between.
#include <algorithm>
#include <memory>
#include <numeric>

#include "primitives.hpp"
#include "spgemm_internal.hpp"

namespace spops {

int64_t g_task_products = 0;
int g_wide_cursors = 2;
int64_t g_heavy_products = kHeavyProducts;

namespace {

inline unsigned nblk(int64_t n, int per) { return grid_blocks(n, per); }

__global__ __launch_bounds__(kBlock) void k_narrow(const int64_t *__restrict__ in, int64_t n, int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}

// ---------------------------------------------------------------------------------------------
// cluster build
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_row_cluster(const int64_t *__restrict__ offsets, int64_t nc,
                                                        int32_t *__restrict__ cid, int32_t *__restrict__ lrow,
                                                        int32_t *__restrict__ csz) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= nc) return;
  const int64_t a = offsets[c], b = offsets[c + 1];
  csz[c] = (int32_t)(b - a);
  for (int64_t t = a; t < b; ++t) { cid[t] = (int32_t)c; lrow[t] = (int32_t)(t - a); }
}

__global__ __launch_bounds__(kBlock) void k_reordered_len(const int64_t *__restrict__ rowptr, const int64_t *__restrict__ perm,
                                                          int64_t n, int64_t *__restrict__ len) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int64_t r = perm[t];
  len[t] = rowptr[r + 1] - rowptr[r];
}

// one wave per reordered row: key = cluster << 32 | col, payload = local row << 40 | entry
__global__ __launch_bounds__(kBlock) void k_cluster_keys(const int64_t *__restrict__ rowptr, const int64_t *__restrict__ col,
                                                         const int64_t *__restrict__ perm, const int64_t *__restrict__ rptr,
                                                         const int32_t *__restrict__ cid, const int32_t *__restrict__ lrow,
                                                         int64_t n, uint64_t *__restrict__ keys, uint64_t *__restrict__ pay) {
  const int64_t t = (int64_t)__builtin_amdgcn_readfirstlane((int)blockIdx.x) * (kBlock / kWave) + wave_id();
  if (t >= n) return;
  const int64_t r = perm[t], p0 = rowptr[r], p1 = rowptr[r + 1], o = rptr[t];
  const uint64_t c = (uint64_t)cid[t], l = (uint64_t)lrow[t];
  for (int64_t j = p0 + lane_id(); j < p1; j += kWave) {
    keys[o + (j - p0)] = (c << 32) | (uint64_t)col[j];
    pay[o + (j - p0)] = (l << 40) | (uint64_t)j;
  }
}

__global__ __launch_bounds__(kBlock) void k_key_heads(const uint64_t *__restrict__ keys, int64_t m, int64_t *__restrict__ head,
                                                      int64_t *__restrict__ ccount) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  const bool h = (a == 0 || keys[a] != keys[a - 1]);
  head[a] = h ? 1 : 0;
  if (h) atomicAdd((unsigned long long *)&ccount[keys[a] >> 32], 1ull);
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_cluster_fill(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ pay,
                                                         const int64_t *__restrict__ uidx, int64_t m,
                                                         const int64_t *__restrict__ crp, const int64_t *__restrict__ crpv,
                                                         const int32_t *__restrict__ csz, const V *__restrict__ aval,
                                                         int32_t *__restrict__ colids, V *__restrict__ values) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  const uint64_t key = keys[a];
  const int64_t c = (int64_t)(key >> 32);
  const int64_t u = uidx[a] + ((a == 0 || key != keys[a - 1]) ? 0 : -1);   // unique index of this key
  const int64_t t = u - crp[c];
  const uint64_t p = pay[a];
  const int64_t l = (int64_t)(p >> 40), j = (int64_t)(p & ((1ull << 40) - 1));
  if (a == 0 || key != keys[a - 1]) colids[u] = (int32_t)(key & 0xffffffffull);
  // a later entry of the same row and column overwrites (the reference's map assignment)
  if (a + 1 < m && keys[a + 1] == key && (int64_t)(pay[a + 1] >> 40) == l) return;
  values[crpv[c] + t * csz[c] + l] = aval[j];
}

__global__ __launch_bounds__(kBlock) void k_mul_csz(const int64_t *__restrict__ nz, const int32_t *__restrict__ csz, int64_t n,
                                                    int64_t *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c < n) out[c] = nz[c] * (int64_t)csz[c];
}

// ---------------------------------------------------------------------------------------------
// SpGEMM: work estimate
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cluster_flop(const int64_t *__restrict__ crp, const int32_t *__restrict__ colids,
                                                         const int64_t *__restrict__ brp, int64_t nc,
                                                         int64_t *__restrict__ flop) {
  const int64_t c = (int64_t)__builtin_amdgcn_readfirstlane((int)blockIdx.x) * (kBlock / kWave) + wave_id();
  if (c >= nc) return;
  int64_t f = 0;
  for (int64_t t = crp[c] + lane_id(); t < crp[c + 1]; t += kWave) {
    const int32_t b = colids[t];
    f += brp[b + 1] - brp[b];
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) f += __shfl_xor(f, m, kWave);
  if (lane_id() == 0) flop[c] = f;
}

// reference flop count (src/spgemm.cpp:24-42): 2 * sum over A entries of nnz(B row)
__global__ __launch_bounds__(kBlock) void k_ref_flop(const int64_t *__restrict__ acol, int64_t annz, int64_t b_rows,
                                                     const int64_t *__restrict__ brp, unsigned long long *__restrict__ out) {
  __shared__ unsigned long long part[kBlock / kWave];
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long f = 0;
  if (j < annz) {
    const int64_t a = acol[j];
    if (a < b_rows) f = (unsigned long long)(brp[a + 1] - brp[a]);
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) f += __shfl_xor(f, m, kWave);
  if (lane_id() == 0) part[wave_id()] = f;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

// ---------------------------------------------------------------------------------------------
// small clusters: one workgroup each, everything in LDS
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hslot(int32_t key, uint32_t mask) { return ((uint32_t)key * 107u) & mask; }

// insert key; returns true when newly inserted
__device__ __forceinline__ bool lds_insert(int32_t *tab, uint32_t mask, int32_t key) {
  uint32_t h = hslot(key, mask);
  for (;;) {
    int32_t old = atomicCAS(&tab[h], -1, key);
    if (old == -1) return true;
    if (old == key) return false;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ uint32_t lds_find(const int32_t *tab, uint32_t mask, int32_t key) {
  uint32_t h = hslot(key, mask);
  while (tab[h] != key) h = (h + 1) & mask;
  return h;
}

// Walk all (t, k) items of cluster c: t over the cluster's columns, k over B row colids[t].
// Waves take columns t in turn; lanes take entries of the B row.
template <typename F>
__device__ __forceinline__ void for_each_product(int64_t t0, int64_t t1, const int32_t *__restrict__ colids,
                                                 const int64_t *__restrict__ brp, F &&f) {
  for (int64_t t = t0 + wave_id(); t < t1; t += kBlock / kWave) {
    const int32_t b = colids[t];
    const int64_t k0 = brp[b], k1 = brp[b + 1];
    for (int64_t k = k0 + lane_id(); k < k1; k += kWave) f(t, k);
  }
}

// symbolic for small clusters: count distinct columns, min and max
__global__ __launch_bounds__(kBlock) void k_sym_small(const int32_t *__restrict__ list, int64_t nlist,
                                                      const int64_t *__restrict__ arp, const int32_t *__restrict__ acol,
                                                      const int64_t *__restrict__ brp, const int32_t *__restrict__ bcol,
                                                      int64_t *__restrict__ nz, int64_t *__restrict__ cmin,
                                                      int64_t *__restrict__ cmax) {
  __shared__ int32_t tab[kSmallTable];
  __shared__ int cnt, mn, mx;
  const int64_t c = list[blockIdx.x];
  for (int i = threadIdx.x; i < kSmallTable; i += kBlock) tab[i] = -1;
  if (threadIdx.x == 0) { cnt = 0; mn = INT32_MAX; mx = -1; }
  __syncthreads();
  for_each_product(arp[c], arp[c + 1], acol, brp, [&](int64_t, int64_t k) {
    const int32_t key = bcol[k];
    if (lds_insert(tab, kSmallTable - 1, key)) {
      atomicAdd(&cnt, 1);
      atomicMin(&mn, key);
      atomicMax(&mx, key);
    }
  });
  __syncthreads();
  if (threadIdx.x == 0) {
    nz[c] = cnt;
    cmin[c] = cnt ? mn : 0;
    cmax[c] = cnt ? mx : 0;
  }
}

// numeric for small clusters: sorted columns + LDS accumulation of the csz-wide blocks
template <typename V>
__global__ __launch_bounds__(kBlock) void k_num_small(const int32_t *__restrict__ list, int64_t nlist,
                                                      const int64_t *__restrict__ arp, const int64_t *__restrict__ arpv,
                                                      const int32_t *__restrict__ csz_a, const int32_t *__restrict__ acol,
                                                      const V *__restrict__ aval, const int64_t *__restrict__ brp,
                                                      const int32_t *__restrict__ bcol, const V *__restrict__ bval,
                                                      const int64_t *__restrict__ crp, const int64_t *__restrict__ crpv,
                                                      int32_t *__restrict__ ccol, V *__restrict__ cval) {
  __shared__ int32_t tab[kSmallTable];
  __shared__ int32_t rank[kSmallTable];
  __shared__ int32_t keys[kSmallTable];
  __shared__ V acc[kSmallAcc];
  __shared__ int cnt;
  const int64_t c = list[blockIdx.x];
  const int csz = csz_a[c];
  const int64_t nzc = crp[c + 1] - crp[c];
  for (int i = threadIdx.x; i < kSmallTable; i += kBlock) { tab[i] = -1; keys[i] = INT32_MAX; }
  for (int i = threadIdx.x; i < kSmallAcc; i += kBlock) acc[i] = V(0);
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for_each_product(arp[c], arp[c + 1], acol, brp, [&](int64_t, int64_t k) {
    const int32_t key = bcol[k];
    if (lds_insert(tab, kSmallTable - 1, key)) keys[atomicAdd(&cnt, 1)] = key;
  });
  __syncthreads();
  // bitonic sort of keys[0 .. P) with P = next power of two >= cnt (padding INT32_MAX)
  const int n = cnt;
  int P = 1;
  while (P < n) P <<= 1;
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < P; i += kBlock) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const int32_t a = keys[i], b = keys[j];
          if ((a > b) == up) { keys[i] = b; keys[j] = a; }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += kBlock) {
    rank[lds_find(tab, kSmallTable - 1, keys[i])] = i;
    ccol[crp[c] + i] = keys[i];
  }
  __syncthreads();
  const int64_t a0 = arp[c], av0 = arpv[c];
  for_each_product(a0, arp[c + 1], acol, brp, [&](int64_t t, int64_t k) {
    const int r = rank[lds_find(tab, kSmallTable - 1, bcol[k])];
    const V b = bval[k];
    const V *a = aval + av0 + (t - a0) * csz;
    for (int l = 0; l < csz; ++l) {
      const V av = a[l];
      if (av != V(0)) atomicAdd(&acc[r * csz + l], av * b);
    }
  });
  __syncthreads();
  const int64_t o = crpv[c];
  for (int64_t i = threadIdx.x; i < nzc * csz; i += kBlock) cval[o + i] = acc[i];
}

// ---------------------------------------------------------------------------------------------
// large clusters: one workgroup each, everything in LDS.
//   symbolic: an LDS bitmap of kLargeTileBits columns per pass over the cluster's B rows
//             gives the count and the min / max column; a second kernel (emit) builds the
//             same bitmaps over [min, max] and writes C's sorted columns (word popcount
//             prefix = rank).
//   numeric:  C's columns of the cluster are taken in chunks of m <= large_acc_entries / csz;
//             a chunk's columns go into an LDS hash (column -> rank inside the chunk) and
//             its csz-wide values accumulate in LDS; each B row contributes only its entries
//             inside the chunk's column range (binary search, B rows are sorted).
// ---------------------------------------------------------------------------------------------

// threads of the bitmap kernels (symbolic, emit, heavy parts): 36 KiB of LDS each, so four
// workgroups per CU; eight waves each keep eight per SIMD in flight
constexpr int kBitmapBlock = 512;

// [k0, k1) narrowed to the entries of a sorted B row with columns in [lo, hi]
__device__ __forceinline__ void row_range(const int32_t *__restrict__ bcol, int64_t &k0, int64_t &k1, int32_t lo,
                                          int32_t hi) {
  int64_t a = k0, b = k1;
  while (a < b) {                            // first column >= lo
    const int64_t m = (a + b) >> 1;
    if (bcol[m] < lo) a = m + 1; else b = m;
  }
  int64_t c = a, d = k1;
  while (c < d) {                            // first column > hi
    const int64_t m = (c + d) >> 1;
    if (bcol[m] <= hi) c = m + 1; else d = m;
  }
  k0 = a;
  k1 = c;
}

// set the bits of the cluster's product columns inside [lo, lo + kLargeTileBits) (all of
// B's columns when whole_rows)
__device__ __forceinline__ void tile_bits(uint32_t *bm, int64_t t0, int64_t t1, const int32_t *__restrict__ acol,
                                          const int64_t *__restrict__ brp, const int32_t *__restrict__ bcol, int32_t lo,
                                          int32_t hi, bool whole_rows) {
  // groups of G threads, one A column (B row) each: G = 16, or more when the cluster has
  // fewer than 16 columns (up to the whole workgroup for one column), so that long B rows of
  // a narrow cluster are read by many lanes.  A row of at most 64 entries is scanned whole
  // with a range check, a longer one is cut to [lo, hi] by binary search first.  Each thread
  // issues four independent loads per step.
  const int64_t na = t1 - t0;
  int G = 16;
  while (G < kBitmapBlock && (int64_t)(kBitmapBlock / G) > 2 * na) G <<= 1;
  const int NG = kBitmapBlock / G;
  const int grp = (int)threadIdx.x / G, sub = (int)threadIdx.x % G;
  constexpr int U = 4;
  for (int64_t t = t0 + grp; t < t1; t += NG) {
    const int32_t b = acol[t];
    int64_t k0 = brp[b], k1 = brp[b + 1];
    if (!whole_rows && k1 - k0 > 64) row_range(bcol, k0, k1, lo, hi);
    for (int64_t k = k0 + sub; k < k1; k += (int64_t)U * G) {
      int32_t col[U];
#pragma unroll
      for (int u = 0; u < U; ++u) col[u] = k + u * G < k1 ? bcol[k + u * G] : -1;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (col[u] >= 0 && (whole_rows || (col[u] >= lo && col[u] <= hi))) {
          const int32_t off = col[u] - lo;
          atomicOr(&bm[off >> 5], 1u << (off & 31));
        }
    }
  }
}

// tile_bits with a cursor per A column (cur[t - t0], in LDS): the entries of a B row below the
// tile were consumed by the earlier tiles (tiles ascend, and the first tile starts at or below
// every column of the cluster's rows), so a row's part of the tile starts at its cursor and
// ends at its first column past hi: found by the scan itself, with no binary search.  Lane
// groups are at most a wave (the end is a group-wide minimum); four loads per lane in flight.
// The end of each row (cend, LDS) is read once with the first cursor, so a row's visit of a
// tile costs one round trip (its entries) instead of three (A column, row end, entries).
// 1020 cursors: 32 KiB of bitmap + 2 x 4080 bytes of cursors and ends leave four workgroups
// per CU.
constexpr int kTileCursors = 1020;
__device__ __forceinline__ void tile_bits_cur(uint32_t *bm, int64_t t0, int64_t t1, const int32_t *__restrict__ bcol,
                                              int32_t lo, int32_t hi, int32_t *cur, const int32_t *cend) {
  const int64_t na = t1 - t0;
  int G = 16;
  while (G < kWave && (int64_t)(kBitmapBlock / G) > 2 * na) G <<= 1;
  const int NG = kBitmapBlock / G;
  const int grp = (int)threadIdx.x / G, sub = (int)threadIdx.x % G;
  constexpr int U = 16;
  for (int64_t t = t0 + grp; t < t1; t += NG) {
    const int32_t k1 = cend[t - t0];
    int32_t k = cur[t - t0];
    bool going = k < k1;                      // uniform inside the group
    while (going) {
      int32_t col[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t pos = k + sub + u * G;
        col[u] = pos < k1 ? bcol[pos] : INT32_MAX;
      }
      int32_t past = INT32_MAX;               // this lane's first position past the tile
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t pos = k + sub + u * G;
        if (col[u] <= hi) {
          if (col[u] >= lo) {
            const int32_t off = col[u] - lo;
            atomicOr(&bm[off >> 5], 1u << (off & 31));
          }
        } else {
          past = min(past, pos < k1 ? pos : k1);
        }
      }
      for (int d = 1; d < G; d <<= 1) past = min(past, __shfl_xor(past, d, kWave));
      if (past != INT32_MAX) {
        k = past;
        going = false;
      } else {
        k += U * G;
        going = k < k1;
        if (!going) k = k1;
      }
    }
    if (sub == 0) cur[t - t0] = k;
  }
}

// writes the columns base + 32 w + b of the set bits b of words[0, nwords) to ccol[out...] in
// ascending order (out advances): rounds of BLK consecutive words, one per thread, with a
// block scan of the per-thread counts; the columns of neighbouring threads are neighbours in
// C, so each store instruction of the wave covers a short range
constexpr int kTileWords = kLargeTileBits / 32;
template <int BLK>
__device__ __forceinline__ void emit_words(const uint32_t *words, int nwords, int32_t base, int64_t &out,
                                           int32_t *__restrict__ ccol, int32_t *wsum) {
  for (int w0 = 0; w0 < nwords; w0 += BLK) {
    const int w = w0 + (int)threadIdx.x;
    uint32_t bits = w < nwords ? words[w] : 0u;
    const int32_t mine = __popc(bits);
    int32_t inc = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int32_t o = __shfl_up(inc, d, kWave);
      if (lane_id() >= d) inc += o;
    }
    if (lane_id() == kWave - 1) wsum[wave_id()] = inc;
    __syncthreads();
    int32_t before = 0, total = 0;
    for (int q = 0; q < BLK / kWave; ++q) {
      if (q < wave_id()) before += wsum[q];
      total += wsum[q];
    }
    int64_t r = out + before + inc - mine;
    while (bits) {
      const int b = __ffs(bits) - 1;
      ccol[r++] = base + w * 32 + b;
      bits &= bits - 1;
    }
    out += total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBitmapBlock) void k_sym_large(const int32_t *__restrict__ list, const int64_t *__restrict__ arp,
                                                      const int32_t *__restrict__ acol, const int64_t *__restrict__ brp,
                                                      const int32_t *__restrict__ bcol, int32_t b_cols,
                                                      int64_t *__restrict__ nz, int64_t *__restrict__ cmin,
                                                      int64_t *__restrict__ cmax, int b32, int32_t *gcur,
                                                      int32_t *gend) {
  __shared__ uint32_t bm[kLargeTileBits / 32];
  __shared__ int32_t cur[kTileCursors], cend[kTileCursors];
  __shared__ int cnt, mn, mx;
  const int64_t c = list[blockIdx.x];
  const int64_t t0 = arp[c], t1 = arp[c + 1];
  if (threadIdx.x == 0) { cnt = 0; mn = INT32_MAX; mx = -1; }
  const bool whole = b_cols <= kLargeTileBits;
  // cursors in LDS, or (clusters of more columns) in device memory at the cluster's columns
  const bool lds_cur = t1 - t0 <= kTileCursors;
  const bool cursors = !whole && b32 && (lds_cur || gcur != nullptr);
  int32_t *cu = lds_cur ? cur : gcur + t0;
  int32_t *ce = lds_cur ? cend : gend + t0;
  if (cursors)
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBitmapBlock) {
      const int32_t b = acol[t];
      cu[t - t0] = (int32_t)brp[b];
      ce[t - t0] = (int32_t)brp[b + 1];
    }
  for (int32_t lo = 0; lo < b_cols; lo += kLargeTileBits) {
    const int32_t hi = (int32_t)min((int64_t)lo + kLargeTileBits, (int64_t)b_cols) - 1;
    for (int w = threadIdx.x; w < kLargeTileBits / 32; w += kBitmapBlock) bm[w] = 0u;
    __syncthreads();
    if (cursors) tile_bits_cur(bm, t0, t1, bcol, lo, hi, cu, ce);
    else tile_bits(bm, t0, t1, acol, brp, bcol, lo, hi, whole);
    __syncthreads();
    int n = 0, first = INT32_MAX, last = -1;
    for (int w = threadIdx.x; w < kLargeTileBits / 32; w += kBitmapBlock) {
      const uint32_t x = bm[w];
      if (x) {
        n += __popc(x);
        first = min(first, lo + w * 32 + __ffs(x) - 1);
        last = max(last, lo + w * 32 + 31 - __clz(x));
      }
    }
    if (n) {
      atomicAdd(&cnt, n);
      atomicMin(&mn, first);
      atomicMax(&mx, last);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nz[c] = cnt;
    cmin[c] = cnt ? mn : 0;
    cmax[c] = cnt ? mx : 0;
  }
}

// C's sorted columns of each large cluster
__global__ __launch_bounds__(kBitmapBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_emit_large(const int32_t *__restrict__ list, const int64_t *__restrict__ arp,
                                                       const int32_t *__restrict__ acol, const int64_t *__restrict__ brp,
                                                       const int32_t *__restrict__ bcol, const int64_t *__restrict__ crp,
                                                       const int64_t *__restrict__ cmin, const int64_t *__restrict__ cmax,
                                                       int32_t *__restrict__ ccol, int b32, int32_t *gcur,
                                                       int32_t *gend) {
  __shared__ uint32_t bm[kLargeTileBits / 32];
  __shared__ int32_t cur[kTileCursors], cend[kTileCursors];
  __shared__ int32_t wsum[kBitmapBlock / kWave];
  const int64_t c = list[blockIdx.x];
  const int64_t t0 = arp[c], t1 = arp[c + 1];
  const int32_t cmn = (int32_t)cmin[c], cmx = (int32_t)cmax[c];
  if (crp[c + 1] == crp[c]) return;
  int64_t out = crp[c];
  const bool lds_cur = t1 - t0 <= kTileCursors;
  const bool cursors = cmx - cmn >= kLargeTileBits && b32 && (lds_cur || gcur != nullptr);
  int32_t *cu = lds_cur ? cur : gcur + t0;
  int32_t *ce = lds_cur ? cend : gend + t0;
  if (cursors)
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBitmapBlock) {
      const int32_t b = acol[t];
      cu[t - t0] = (int32_t)brp[b];
      ce[t - t0] = (int32_t)brp[b + 1];
    }
  for (int32_t lo = cmn; lo <= cmx; lo += kLargeTileBits) {
    const int32_t hi = (int32_t)min((int64_t)lo + kLargeTileBits - 1, (int64_t)cmx);
    const int nwords = ((hi - lo) >> 5) + 1;          // words that can hold bits of [lo, hi]
    for (int w = threadIdx.x; w < nwords; w += kBitmapBlock) bm[w] = 0u;
    __syncthreads();
    if (cursors) tile_bits_cur(bm, t0, t1, bcol, lo, hi, cu, ce);
    else tile_bits(bm, t0, t1, acol, brp, bcol, lo, hi, cmx - cmn < kLargeTileBits);
    __syncthreads();
    emit_words<kBitmapBlock>(bm, nwords, lo, out, ccol, wsum);
  }
}

// clusters of many products ("heavy"): the symbolic bitmap over all of B's columns is built
// by several workgroups, each taking a part of the cluster's A columns (at most kTileCursors,
// about kHeavyPartProducts products) and writing its bitmap to a slice of its own; a merge pass
// ORs a cluster's slices into the bitmap kept by the plan and counts it (nz, min, max column),
// and the emit pass writes C's columns straight from that bitmap.
struct HeavyPart {
  int64_t t0, t1;          // A columns of the part (cluster column pointers)
  int32_t slice, pad;      // bitmap slice written
};

__global__ __launch_bounds__(kBitmapBlock) void k_heavy_part(const HeavyPart *__restrict__ parts,
                                                       const int32_t *__restrict__ acol, const int64_t *__restrict__ brp,
                                                       const int32_t *__restrict__ bcol, int32_t b_cols, int64_t words,
                                                       uint32_t *__restrict__ slices, int b32) {
  __shared__ uint32_t bm[kLargeTileBits / 32];
  __shared__ int32_t cur[kTileCursors], cend[kTileCursors];
  const HeavyPart p = parts[blockIdx.x];
  const bool whole = b_cols <= kLargeTileBits;
  const bool cursors = !whole && b32 && p.t1 - p.t0 <= kTileCursors;
  if (cursors)
    for (int64_t t = p.t0 + threadIdx.x; t < p.t1; t += kBitmapBlock) {
      const int32_t b = acol[t];
      cur[t - p.t0] = (int32_t)brp[b];
      cend[t - p.t0] = (int32_t)brp[b + 1];
    }
  uint32_t *out = slices + (int64_t)p.slice * words;
  for (int64_t lo64 = 0; lo64 < b_cols; lo64 += kLargeTileBits) {
    const int32_t lo = (int32_t)lo64;
    const int32_t hi = (int32_t)(min(lo64 + kLargeTileBits, (int64_t)b_cols) - 1);
    const int nwords = ((hi - lo) >> 5) + 1;
    for (int w = threadIdx.x; w < nwords; w += kBitmapBlock) bm[w] = 0u;
    __syncthreads();
    if (cursors) tile_bits_cur(bm, p.t0, p.t1, bcol, lo, hi, cur, cend);
    else tile_bits(bm, p.t0, p.t1, acol, brp, bcol, lo, hi, whole);
    __syncthreads();
    for (int w = threadIdx.x; w < nwords; w += kBitmapBlock) out[(lo >> 5) + w] = bm[w];
    __syncthreads();
  }
}

// per heavy cluster {cluster, first slice, end slice, index h}: OR of its slices -> bitmap h,
// and the cluster's nz / min / max column
__global__ __launch_bounds__(kBlock) void k_heavy_merge(const int4 *__restrict__ heavy, const uint32_t *__restrict__ slices,
                                                        int64_t words, uint32_t *__restrict__ gbm,
                                                        int64_t *__restrict__ nz, int64_t *__restrict__ cmin,
                                                        int64_t *__restrict__ cmax) {
  __shared__ int64_t scnt[kBlock / kWave];
  __shared__ int64_t smn[kBlock / kWave], smx[kBlock / kWave];
  const int4 hv = heavy[blockIdx.x];
  uint32_t *dst = gbm + (int64_t)hv.w * words;
  int64_t n = 0, first = INT64_MAX, last = -1;
  for (int64_t w = threadIdx.x; w < words; w += kBlock) {
    uint32_t x = 0u;
    for (int q = hv.y; q < hv.z; ++q) x |= slices[(int64_t)q * words + w];
    dst[w] = x;
    if (x) {
      n += __popc(x);
      first = min(first, w * 32 + __ffs(x) - 1);
      last = max(last, w * 32 + 31 - __clz(x));
    }
  }
  for (int d = 1; d < kWave; d <<= 1) {
    n += __shfl_xor(n, d, kWave);
    first = min(first, __shfl_xor(first, d, kWave));
    last = max(last, __shfl_xor(last, d, kWave));
  }
  if (lane_id() == 0) {
    scnt[wave_id()] = n;
    smn[wave_id()] = first;
    smx[wave_id()] = last;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / kWave; ++q) {
      n += scnt[q];
      first = min(first, smn[q]);
      last = max(last, smx[q]);
    }
    nz[hv.x] = n;
    cmin[hv.x] = n ? first : 0;
    cmax[hv.x] = n ? last : 0;
  }
}

// C's sorted columns of heavy clusters {cluster, bitmap h} from the kept bitmaps
__global__ __launch_bounds__(kBlock) void k_heavy_emit(const int2 *__restrict__ list, const uint32_t *__restrict__ gbm,
                                                       int64_t words, const int64_t *__restrict__ crp,
                                                       const int64_t *__restrict__ cmin, const int64_t *__restrict__ cmax,
                                                       int32_t *__restrict__ ccol) {
  __shared__ int32_t wsum[kBlock / kWave];
  const int2 e = list[blockIdx.x];
  const int64_t c = e.x;
  if (crp[c + 1] == crp[c]) return;
  int64_t out = crp[c];
  const int64_t w0 = cmin[c] >> 5, w1 = cmax[c] >> 5;
  const uint32_t *src = gbm + (int64_t)e.y * words + w0;
  for (int64_t w = 0; w <= w1 - w0; w += kTileWords) {
    const int nwords = (int)min((int64_t)kTileWords, w1 - w0 + 1 - w);
    emit_words<kBlock>(src + w, nwords, (int32_t)((w0 + w) * 32), out, ccol, wsum);
  }
}

// CR: A values of a cluster column held in registers per pass over its B row entries (4 for
// clusters of at most 4 rows, kMaxClusterRows otherwise); the products of an entry stop at
// the cluster's size.  With B row cursors, lane group g of a wave keeps, for the first row
// it handles (all of its rows when one round of the workgroup covers the cluster), that
// row's next GL entries and, when CR <= 4, its A values in registers from one chunk to the
// next: the entries are loaded when the row leaves a chunk, so the next chunk starts without
// a round trip to memory.
// A task is {cluster, first chunk, end chunk}: a cluster with much work is split over several
// workgroups by ranges of its chunks (each task's row cursors start at its first chunk's
// column, one binary search per row), so that one cluster does not hold the kernel's tail.
// CURS: A columns whose cursors fit LDS (kLargeCursors, or kLargeCursorsWide for a variant
// with 78 KB of LDS, two workgroups per CU, that takes the tasks of clusters with more
// columns: cursors instead of two binary searches per row and chunk).
template <typename V, int CR, int CURS>
__global__ __launch_bounds__(kLargeBlock) __attribute__((amdgpu_waves_per_eu(CR <= 4 && CURS <= kLargeCursors ? 8 : 1)))
void k_num_large(const int4 *__restrict__ tasks, const int64_t *__restrict__ arp,
                                                           const int64_t *__restrict__ arpv, const int32_t *__restrict__ csz_a,
                                                           const int32_t *__restrict__ acol, const V *__restrict__ aval,
                                                           const int64_t *__restrict__ brp, const int32_t *__restrict__ bcol,
                                                           const V *__restrict__ bval, const int64_t *__restrict__ crp,
                                                           const int64_t *__restrict__ crpv, const int32_t *__restrict__ ccol,
                                                           V *__restrict__ cval, bool b32) {
  constexpr int HS = large_hash_slots(CR);
  __shared__ V acc[large_acc_entries(CR)];
  __shared__ int32_t hkey[HS];
  __shared__ int16_t hval[HS];
  // per A column (when the cluster has at most kLargeCursors and B's offsets fit 32 bits):
  // the next unread entry of its B row, the row's end, and the column at that entry
  // (INT32_MAX: the row is done), so a chunk reads only B entries and skips rows at once
  __shared__ int32_t rpos[CURS];
  __shared__ int32_t rend[CURS];
  __shared__ int32_t nextcol[CURS];
  __shared__ int next_row;                     // next row (relative to the cluster's first) to claim
  constexpr int W = kLargeBlock / kWave;
  const int4 tk = tasks[blockIdx.x];
  const int64_t c = tk.x;
  const int csz = csz_a[c];
  const int64_t a0 = arp[c], a1 = arp[c + 1], av0 = arpv[c];
  const int64_t c0 = crp[c], cv0 = crpv[c];
  const int M = large_chunk_cols(csz, CR);
  const int64_t jbeg = (int64_t)tk.y * M, jend = min((int64_t)tk.z * M, crp[c + 1] - c0);
  const bool cursors = a1 - a0 <= CURS && b32;
  if (cursors) {
    const int32_t lo0 = jbeg > 0 && jbeg < jend ? ccol[c0 + jbeg] : INT32_MIN;
    for (int64_t t = a0 + threadIdx.x; t < a1; t += kLargeBlock) {
      const int32_t b = acol[t];
      int32_t kb = (int32_t)brp[b];
      const int32_t ke = (int32_t)brp[b + 1];
      if (lo0 != INT32_MIN) {                 // first entry of the row at or past the task's first column
        int32_t hi_ = ke;
        while (kb < hi_) {
          const int32_t mid = kb + ((hi_ - kb) >> 1);
          if (bcol[mid] < lo0) kb = mid + 1; else hi_ = mid;
        }
      }
      rpos[t - a0] = kb;
      rend[t - a0] = ke;
      nextcol[t - a0] = kb < ke ? bcol[kb] : INT32_MAX;
    }
  }
  const int lane = lane_id();
  // the prefetch slot (see above): entries from position pf_k (-1: none), A values once loaded
  constexpr bool kCacheA = CR <= 4;           // A values kept from chunk to chunk (registers)
  constexpr int CA = kCacheA ? CR : 1;
  int32_t pf_key = INT32_MAX, pf_k = -1;
  V pf_bv = V(0);
  V pf_a[CA];
  bool pf_has_a = false;
#pragma unroll
  for (int l = 0; l < CA; ++l) pf_a[l] = V(0);
  // lanes per B row: a power of two in [kLargeGroupLanes, 64], about 3.2 M / rows (each
  // output column takes ~1.2 B entries, shared by the cluster's rows, so a row brings ~1.2 M /
  // rows entries to a chunk on average).  Each step over a row is a dependent round trip to
  // memory and a chunk lasts as long as its longest row part, so groups are made wider than
  // the average part; narrower groups would let more rows run at once.  Measured factor
  // (3.2 vs 1.2, 2.4, 4.8, 6.4): R-MAT 16/18/20 91.7 -> 96.3, 96.7 -> 104.3, 77.8 -> 85.0 GFLOP/s.
  int GL = kLargeGroupLanes;
  while (GL < kWave && (int64_t)GL * 5 * (a1 - a0) < (int64_t)16 * M) GL <<= 1;
  GL = __builtin_amdgcn_readfirstlane(GL);
  const int G = kWave / GL;
  const int64_t first_base = a0 + (int64_t)wave_id() * G;
  for (int64_t j0 = jbeg; j0 < jend; j0 += M) {
    const int m = (int)min((int64_t)M, jend - j0);
    const int32_t lo = ccol[c0 + j0], hi = ccol[c0 + j0 + m - 1];
    for (int i = threadIdx.x; i < HS; i += kLargeBlock) hkey[i] = -1;
    for (int i = threadIdx.x; i < m * csz; i += kLargeBlock) acc[i] = V(0);
    if (threadIdx.x == 0) next_row = W * G;   // rows past the first round, handed out on demand
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kLargeBlock) {
      const int32_t key = ccol[c0 + j0 + i];
      uint32_t h = hslot(key, HS - 1);
      while (atomicCAS(&hkey[h], -1, key) != -1) h = (h + 1) & (HS - 1);
      hval[h] = (int16_t)i;
    }
    __syncthreads();
    // lane groups of GL lanes, each scanning one B row
    const int grp = lane / GL, sub = lane % GL;
    const uint64_t gmask = ((GL == 64) ? ~0ull : ((1ull << GL) - 1)) << (grp * GL);
    if (CR <= 4 || csz <= CR) {              // (the CR = 4 kernel only gets clusters of <= 4 rows)
      // One pass per row.  Each lane group starts with its row of the first round (its
      // prefetch-slot row) and then claims the next row from the LDS counter as soon as its
      // row is done, so a few long B rows do not hold the other groups: with rows dealt out
      // round-robin the slowest wave of a chunk took twice the mean number of steps.
      int64_t t = first_base + grp;
      bool slot = cursors, first = true, going = false, fetch = false, done = false;
      int64_t k = 0, ke = 0;
      V ar[CR];
#pragma unroll
      for (int l = 0; l < CR; ++l) ar[l] = V(0);
      // row t's entries in this chunk start at k (false: none)
      auto start_row = [&]() {
        bool live = t < a1 && !(cursors && nextcol[t - a0] > hi);
        k = 0;
        ke = 0;
        if (live) {
          if (cursors) {
            k = rpos[t - a0];
            ke = rend[t - a0];
          } else {
            const int32_t b = acol[t];
            k = brp[b];
            ke = brp[b + 1];
            int64_t k1 = ke;
            row_range(bcol, k, k1, lo, hi);
          }
          live = k < ke;
        }
        if (live) {
          if (kCacheA && slot && pf_has_a) {
#pragma unroll
            for (int l = 0; l < CA; ++l) ar[l] = pf_a[l];
          } else {
            const V *a = aval + av0 + (t - a0) * csz;
#pragma unroll
            for (int l = 0; l < CR; ++l) ar[l] = l < csz ? a[l] : V(0);
            if (kCacheA && slot) {
#pragma unroll
              for (int l = 0; l < CA; ++l) pf_a[l] = ar[l];
              pf_has_a = true;
            }
          }
        }
        first = true;
        return live;
      };
      going = start_row();
      fetch = !going;
      for (;;) {
        // groups without a row claim one (one claim per group and iteration)
        int claim = 0;
        if (fetch && !done && sub == 0) claim = atomicAdd(&next_row, 1);
        claim = __shfl(claim, grp * GL, kWave);
        if (fetch && !done) {
          if (a0 + claim >= a1) {
            done = true;
          } else {
            t = a0 + claim;
            slot = false;
            going = start_row();
          }
          fetch = !going && !done;
        }
        if (!__ballot(going)) {
          if (__ballot(fetch)) continue;
          break;
        }
        // NE entries of the row per lane and step (loads in flight together; the entries past
        // the first set count only when the whole first set is inside the chunk, the row being
        // sorted).  NE = 2: R-MAT 16 / 20 95.1 -> 97.9 / 84.4 -> 87.0 GFLOP/s; 3 and 4 spill
        // registers at the 64 VGPRs of eight waves per SIMD (86.0 / 84.5 at R-MAT 20).
        constexpr int NE = 2;
        const bool use_pf = slot && first && pf_k == (int32_t)k;
        int32_t key[NE];
        V bv[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const int64_t kq = k + sub + (int64_t)e * GL;
          key[e] = INT32_MAX;
          bv[e] = V(0);
          if (going && kq < ke) {
            if (e == 0 && use_pf) {
              key[e] = pf_key;
              bv[e] = pf_bv;
            } else {
              key[e] = bcol[kq];
              bv[e] = bval[kq];
            }
          }
        }
        first = false;
        int n = 0, nfirst = -1;
        int32_t first_out = INT32_MAX;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const bool in = going && key[e] <= hi;
          const int ne = __builtin_popcountll(__ballot(in) & gmask);
          const int32_t fo = __shfl(key[e], grp * GL + (ne < GL ? ne : 0), kWave);
          if (nfirst < 0 && ne < GL) { nfirst = e; first_out = fo; }
          n += ne;
          int h = -1;
          if (in) {                            // every such column is one of the chunk's
            uint32_t p = hslot(key[e], HS - 1);
            for (int probe = 0; probe < HS; ++probe, p = (p + 1) & (HS - 1))
              if (hkey[p] == key[e]) { h = (int)p; break; }
          }
          if (h >= 0) {
            V *dst = acc + (int)hval[h] * csz;
#pragma unroll
            for (int l = 0; l < CR; ++l) {
              if (l >= csz) break;
              if (ar[l] != V(0)) atomicAdd(&dst[l], ar[l] * bv[e]);
            }
          }
        }
        if (going) {
          k += n;
          if (n < NE * GL) {                   // the row's part of the chunk is done
            const int32_t nxt = first_out;     // its first column past hi (INT32_MAX: row end)
            going = false;
            fetch = true;
            if (cursors && sub == 0) {
              rpos[t - a0] = (int32_t)k;
              nextcol[t - a0] = nxt;
            }
            if (slot) {                        // the entries it brings to its next chunk
              pf_k = nxt == INT32_MAX ? -1 : (int32_t)k;
              const int64_t k2 = k + sub;
              if (pf_k >= 0 && k2 < ke) {
                pf_key = bcol[k2];
                pf_bv = bval[k2];
              } else {
                pf_key = INT32_MAX;
                pf_bv = V(0);
              }
            }
          }
        }
      }
    } else {
    for (int64_t base = a0 + (int64_t)wave_id() * G; base < a1; base += (int64_t)W * G) {
      // every lane of the wave takes part in the ballots below, groups without a row idle
      const int64_t t = base + grp;
      bool live = t < a1 && !(cursors && nextcol[t - a0] > hi);
      int64_t ke = 0, k = 0;
      const V *a = nullptr;
      if (live) {
        if (cursors) {
          k = rpos[t - a0];
          ke = rend[t - a0];
        } else {
          const int32_t b = acol[t];
          k = brp[b];
          ke = brp[b + 1];
          int64_t k1 = ke;
          row_range(bcol, k, k1, lo, hi);
        }
        live = k < ke;
        a = aval + av0 + (t - a0) * csz;
      }
      const int64_t kstart = k;
      int32_t nxt = INT32_MAX;
      const bool slot = cursors && base == first_base;        // wave-uniform
      for (int l0 = 0; l0 < csz; l0 += CR) {                  // one pass unless csz > CR
        const int lmax = min(csz - l0, CR);
        V ar[CR];
        if (kCacheA && slot && pf_has_a && csz <= CR) {
#pragma unroll
          for (int l = 0; l < CA; ++l) ar[l] = live ? pf_a[l] : V(0);
        } else {
#pragma unroll
          for (int l = 0; l < CR; ++l) ar[l] = (live && l < lmax) ? a[l0 + l] : V(0);
          if (kCacheA && slot && live && csz <= CR) {
#pragma unroll
            for (int l = 0; l < CA; ++l) pf_a[l] = ar[l];
            pf_has_a = true;
          }
        }
        k = kstart;
        bool going = live;
        bool first = true;
        while (__ballot(going)) {
          const int64_t kk = k + sub;
          const bool have = going && kk < ke;
          const bool use_pf = slot && first && pf_k == (int32_t)k;
          int32_t key = INT32_MAX;
          V bv = V(0);
          if (have) {
            if (use_pf) {
              key = pf_key;
              bv = pf_bv;
            } else {
              key = bcol[kk];
              bv = bval[kk];
            }
          }
          first = false;
          const bool in = going && key <= hi;
          const int n = __builtin_popcountll(__ballot(in) & gmask);
          int h = -1;
          if (in) {                            // every such column is one of the chunk's
            uint32_t p = hslot(key, HS - 1);
            for (int probe = 0; probe < HS; ++probe, p = (p + 1) & (HS - 1))
              if (hkey[p] == key) { h = (int)p; break; }
          }
          if (h >= 0) {
            V *dst = acc + (int)hval[h] * csz + l0;
#pragma unroll
            for (int l = 0; l < CR; ++l) {
              if (l >= lmax) break;
              if (ar[l] != V(0)) atomicAdd(&dst[l], ar[l] * bv);
            }
          }
          const int32_t first_out = __shfl(key, grp * GL + (n < GL ? n : 0), kWave);
          if (going) {
            k += n;
            if (n < GL) {
              nxt = first_out;                 // first column past hi (INT32_MAX past the row)
              going = false;
            }
          }
        }
      }
      if (cursors && live && sub == 0) {
        rpos[t - a0] = (int32_t)k;
        nextcol[t - a0] = nxt;
      }
      if (slot && live) {                    // the row's entries for the next chunk it takes part in
        pf_k = nxt == INT32_MAX ? -1 : (int32_t)k;
        const int64_t kk = k + sub;
        if (pf_k >= 0 && kk < ke) {
          pf_key = bcol[kk];
          pf_bv = bval[kk];
        } else {
          pf_key = INT32_MAX;
          pf_bv = V(0);
        }
      }
    }
    }
    __syncthreads();
    V *out = cval + cv0 + j0 * csz;
    for (int i = threadIdx.x; i < m * csz; i += kLargeBlock) out[i] = acc[i];
    __syncthreads();
  }
}

// flag = 1 when some row of B has columns out of ascending order
__global__ __launch_bounds__(kBlock) void k_rows_unsorted(const int64_t *__restrict__ brp, int64_t b_rows,
                                                          const int32_t *__restrict__ bcol, int *__restrict__ flag) {
  const int64_t r = (int64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (r >= b_rows) return;
  const int64_t k0 = brp[r], k1 = brp[r + 1];
  bool bad = false;
  for (int64_t k = k0 + 1 + lane_id(); k < k1; k += kWave) bad |= bcol[k] < bcol[k - 1];
  if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1);
}

__global__ __launch_bounds__(kBlock) void k_row_keys(const int64_t *__restrict__ brp, int64_t b_rows,
                                                     const int32_t *__restrict__ bcol, uint64_t *__restrict__ keys,
                                                     uint64_t *__restrict__ idx) {
  const int64_t r = (int64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (r >= b_rows) return;
  for (int64_t k = brp[r] + lane_id(); k < brp[r + 1]; k += kWave) {
    keys[k] = ((uint64_t)r << 32) | (uint32_t)bcol[k];
    idx[k] = (uint64_t)k;
  }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_row_sorted_fill(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ idx,
                                                            int64_t n, const V *__restrict__ val, int32_t *__restrict__ scol,
                                                            V *__restrict__ sval) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  scol[i] = (int32_t)(uint32_t)keys[i];
  sval[i] = val[idx[i]];
}

// ---------------------------------------------------------------------------------------------
// C (cluster) -> CSR in original row order, dropping exact zeros
// ---------------------------------------------------------------------------------------------
// one wave per (cluster, local row)
template <typename V>
__global__ __launch_bounds__(kBlock) void k_c2csr_count(const int64_t *__restrict__ offsets, const int64_t *__restrict__ perm,
                                                        const int32_t *__restrict__ cid, const int32_t *__restrict__ lrow,
                                                        const int32_t *__restrict__ csz, const int64_t *__restrict__ crp,
                                                        const int64_t *__restrict__ crpv, const V *__restrict__ cval,
                                                        int64_t n, int64_t *__restrict__ cnt) {
  const int64_t t = (int64_t)__builtin_amdgcn_readfirstlane((int)blockIdx.x) * (kBlock / kWave) + wave_id();
  if (t >= n) return;
  const int64_t c = cid[t];
  const int l = lrow[t], s = csz[c];
  int64_t m = 0;
  for (int64_t q = lane_id(); q < crp[c + 1] - crp[c]; q += kWave) m += (cval[crpv[c] + q * s + l] != V(0)) ? 1 : 0;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) m += __shfl_xor(m, d, kWave);
  if (lane_id() == 0) cnt[perm[t]] = m;
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_c2csr_fill(const int64_t *__restrict__ perm, const int32_t *__restrict__ cid,
                                                       const int32_t *__restrict__ lrow, const int32_t *__restrict__ csz,
                                                       const int64_t *__restrict__ crp, const int64_t *__restrict__ crpv,
                                                       const int32_t *__restrict__ ccol, const V *__restrict__ cval, int64_t n,
                                                       const int64_t *__restrict__ orp, int64_t *__restrict__ ocol,
                                                       V *__restrict__ oval) {
  const int64_t t = (int64_t)__builtin_amdgcn_readfirstlane((int)blockIdx.x) * (kBlock / kWave) + wave_id();
  if (t >= n) return;
  const int64_t c = cid[t];
  const int l = lrow[t], s = csz[c];
  const int64_t nzc = crp[c + 1] - crp[c];
  int64_t o = orp[perm[t]];
  const uint64_t lt = lane_id() == 0 ? 0ull : (~0ull >> (64 - lane_id()));
  for (int64_t q0 = 0; q0 < nzc; q0 += kWave) {
    const int64_t q = q0 + lane_id();
    V v = V(0);
    if (q < nzc) v = cval[crpv[c] + q * s + l];
    const bool keep = q < nzc && v != V(0);
    const uint64_t bal = __ballot(keep);
    if (keep) {
      const int64_t p = o + __popcll(bal & lt);
      ocol[p] = ccol[crp[c] + q];
      oval[p] = v;
    }
    o += __popcll(bal);
  }
}

}  // namespace

// =============================================================================================
// host orchestration
// =============================================================================================
void cluster_build(spops_cluster_s &A, int64_t n, int64_t n_cols, int64_t nnz, const int64_t *rowptr, const int64_t *col,
                   const void *values, int vt, const int64_t *h_perm, int64_t nc, const int64_t *h_offsets, hipStream_t s) {
  A.csr_rows = n;
  A.rows = nc;
  A.cols = n_cols;
  A.value_type = vt;
  A.perm.upload(std::vector<int64_t>(h_perm, h_perm + n));
  A.offsets.upload(std::vector<int64_t>(h_offsets, h_offsets + nc + 1));
  A.max_cluster_sz = 0;
  for (int64_t c = 0; c < nc; ++c) A.max_cluster_sz = std::max(A.max_cluster_sz, h_offsets[c + 1] - h_offsets[c]);
  A.cid.alloc((size_t)std::max<int64_t>(n, 1) * 4);
  A.lrow.alloc((size_t)std::max<int64_t>(n, 1) * 4);
  A.cluster_sz.alloc((size_t)std::max<int64_t>(nc, 1) * 4);
  if (nc > 0)
    hipLaunchKernelGGL(k_row_cluster, dim3(nblk(nc, kBlock)), dim3(kBlock), 0, s, A.offsets.as<int64_t>(), nc,
                       A.cid.as<int32_t>(), A.lrow.as<int32_t>(), A.cluster_sz.as<int32_t>());
  // reordered row offsets
  DevBuf len, rptr;
  len.alloc((size_t)std::max<int64_t>(n, 1) * 8);
  rptr.alloc((size_t)std::max<int64_t>(n, 1) * 8);
  if (n > 0) {
    hipLaunchKernelGGL(k_reordered_len, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, rowptr, A.perm.as<int64_t>(), n,
                       len.as<int64_t>());
    SPOPS_HIP_CHECK(hipGetLastError());
    exclusive_scan_i64(len.as<int64_t>(), rptr.as<int64_t>(), n, s, nullptr);
  }
  DevBuf keys, pay, head, uidx, ccount;
  keys.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  pay.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  if (n > 0)
    hipLaunchKernelGGL(k_cluster_keys, dim3(nblk(n, kBlock / kWave)), dim3(kBlock), 0, s, rowptr, col,
                       A.perm.as<int64_t>(), rptr.as<int64_t>(), A.cid.as<int32_t>(), A.lrow.as<int32_t>(), n,
                       keys.as<uint64_t>(), pay.as<uint64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  len.release();
  rptr.release();
  const int cbits = bit_width_u64((uint64_t)std::max<int64_t>(nc - 1, 0));
  radix_sort_u64(keys.as<uint64_t>(), pay.p, 8, nnz, 0, 32 + cbits, s);
  head.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  uidx.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  ccount.alloc((size_t)(nc + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(ccount.p, 0, (size_t)(nc + 1) * 8, s));
  if (nnz > 0)
    hipLaunchKernelGGL(k_key_heads, dim3(nblk(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(), nnz,
                       head.as<int64_t>(), ccount.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  int64_t nnzc = 0;
  exclusive_scan_i64(head.as<int64_t>(), uidx.as<int64_t>(), nnz, s, &nnzc);
  // uidx[a] = number of heads before a; for a head that is its unique index, for a
  // non-head it is one past the index of its key (k_cluster_fill corrects by -1)
  A.rowptr.alloc((size_t)(nc + 1) * 8);
  exclusive_scan_i64(ccount.as<int64_t>(), A.rowptr.as<int64_t>(), nc + 1, s, nullptr);
  DevBuf vcount;
  vcount.alloc((size_t)(nc + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(vcount.p, 0, (size_t)(nc + 1) * 8, s));
  if (nc > 0)
    hipLaunchKernelGGL(k_mul_csz, dim3(nblk(nc, kBlock)), dim3(kBlock), 0, s, ccount.as<int64_t>(),
                       A.cluster_sz.as<int32_t>(), nc, vcount.as<int64_t>());
  A.rowptr_val.alloc((size_t)(nc + 1) * 8);
  int64_t nnzv = 0;
  exclusive_scan_i64(vcount.as<int64_t>(), A.rowptr_val.as<int64_t>(), nc + 1, s, &nnzv);
  A.nnzc = nnzc;
  A.nnzv = nnzv;
  A.colids.alloc((size_t)std::max<int64_t>(nnzc, 1) * 4);
  A.values.alloc((size_t)std::max<int64_t>(nnzv, 1) * value_bytes(vt));
  SPOPS_HIP_CHECK(hipMemsetAsync(A.values.p, 0, (size_t)std::max<int64_t>(nnzv, 1) * value_bytes(vt), s));
  if (nnz > 0) {
    if (vt == SPOPS_VALUE_F64)
      hipLaunchKernelGGL(k_cluster_fill<double>, dim3(nblk(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(),
                         pay.as<uint64_t>(), uidx.as<int64_t>(), nnz, A.rowptr.as<int64_t>(), A.rowptr_val.as<int64_t>(),
                         A.cluster_sz.as<int32_t>(), static_cast<const double *>(values), A.colids.as<int32_t>(),
                         A.values.as<double>());
    else
      hipLaunchKernelGGL(k_cluster_fill<float>, dim3(nblk(nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(),
                         pay.as<uint64_t>(), uidx.as<int64_t>(), nnz, A.rowptr.as<int64_t>(), A.rowptr_val.as<int64_t>(),
                         A.cluster_sz.as<int32_t>(), static_cast<const float *>(values), A.colids.as<int32_t>(),
                         A.values.as<float>());
    SPOPS_HIP_CHECK(hipGetLastError());
  }
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
}

namespace {

// ids by decreasing flop[id], ties in their current order: a stable LSD radix sort of
// (max - flop) with 16-bit digits (a comparison sort of the ~10^5..10^6 large clusters took
// tens of milliseconds on the host)
void sort_by_work_desc(std::vector<int32_t> &ids, const std::vector<int64_t> &flop) {
  const size_t n = ids.size();
  if (n < 2) return;
  int64_t mx = 0;
  for (int32_t c : ids) mx = std::max(mx, flop[c]);
  std::vector<uint64_t> key(n), ktmp(n);
  std::vector<int32_t> tmp(n);
  for (size_t i = 0; i < n; ++i) key[i] = (uint64_t)(mx - flop[ids[i]]);
  std::vector<size_t> cnt(65536 + 1);
  for (int sh = 0; sh < 64 && ((uint64_t)mx >> sh) != 0; sh += 16) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) ++cnt[((key[i] >> sh) & 0xffffu) + 1];
    for (size_t d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t at = cnt[(key[i] >> sh) & 0xffffu]++;
      tmp[at] = ids[i];
      ktmp[at] = key[i];
    }
    ids.swap(tmp);
    key.swap(ktmp);
  }
}

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    SPOPS_HIP_CHECK(hipEventCreate(&a));
    SPOPS_HIP_CHECK(hipEventCreate(&b));
    SPOPS_HIP_CHECK(hipEventRecord(a, s));
  }
  double stop() {
    SPOPS_HIP_CHECK(hipEventRecord(b, s));
    SPOPS_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    SPOPS_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    SPOPS_HIP_CHECK(hipEventRecord(a, s));
    return ms;
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

}  // namespace

// ---- dense accumulation for clusters with very many products (R-MAT hubs: tens of
// thousands of A columns, C rows covering most columns).  The cluster's rows of C accumulate
// in a dense block D[l * range + (col - cmin)] (l < csz, range = cmax - cmin + 1) in device
// memory: one workgroup per kDenseTaskCols consecutive A columns of a cluster, 16 lanes per B
// row, atomic adds; then the block is gathered at the cluster's sorted output columns.
struct DenseTask {
  int64_t cluster, t0, t1, off;   // A columns [t0, t1) of the cluster, D at scratch + off
};

template <typename V>
__global__ __launch_bounds__(kBlock) void k_dense_products(const DenseTask *__restrict__ tasks,
                                                           const int64_t *__restrict__ arp,
                                                           const int64_t *__restrict__ arpv,
                                                           const int32_t *__restrict__ csz_a,
                                                           const int32_t *__restrict__ acol, const V *__restrict__ aval,
                                                           const int64_t *__restrict__ brp,
                                                           const int32_t *__restrict__ bcol, const V *__restrict__ bval,
                                                           const int64_t *__restrict__ cmin,
                                                           const int64_t *__restrict__ cmax, V *__restrict__ scratch) {
  const DenseTask tk = tasks[blockIdx.x];
  const int64_t c = tk.cluster;
  const int csz = csz_a[c];
  const int64_t lo = cmin[c], range = cmax[c] - lo + 1;
  V *D = scratch + tk.off;
  const int64_t a0 = arp[c], av0 = arpv[c];
  constexpr int GL = 16, GPB = kBlock / GL;
  const int g = threadIdx.x / GL, sub = threadIdx.x % GL;
  for (int64_t t = tk.t0 + g; t < tk.t1; t += GPB) {
    const int32_t b = acol[t];
    const V *a = aval + av0 + (t - a0) * csz;
    const int64_t k0 = brp[b], k1 = brp[b + 1];
    for (int l0 = 0; l0 < csz; l0 += kMaxClusterRows) {
      V ar[kMaxClusterRows];
#pragma unroll
      for (int l = 0; l < kMaxClusterRows; ++l) ar[l] = l0 + l < csz ? a[l0 + l] : V(0);
      for (int64_t k = k0 + sub; k < k1; k += GL) {
        const int64_t col = (int64_t)bcol[k] - lo;
        const V bv = bval[k];
#pragma unroll
        for (int l = 0; l < kMaxClusterRows; ++l)
          if (ar[l] != V(0)) atomicAdd(&D[(int64_t)(l0 + l) * range + col], ar[l] * bv);
      }
    }
  }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_dense_gather(const int32_t *__restrict__ list,
                                                         const int64_t *__restrict__ offs,
                                                         const int32_t *__restrict__ csz_a,
                                                         const int64_t *__restrict__ cmin,
                                                         const int64_t *__restrict__ cmax, const V *__restrict__ scratch,
                                                         const int64_t *__restrict__ crp,
                                                         const int64_t *__restrict__ crpv,
                                                         const int32_t *__restrict__ ccol, V *__restrict__ cval) {
  const int64_t c = list[blockIdx.y];
  const int csz = csz_a[c];
  const int64_t lo = cmin[c], range = cmax[c] - lo + 1;
  const V *D = scratch + offs[blockIdx.y];
  const int64_t j0 = crp[c], nzc = crp[c + 1] - j0, cv0 = crpv[c];
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nzc; j += (int64_t)gridDim.x * kBlock) {
    const int64_t col = (int64_t)ccol[j0 + j] - lo;
    for (int l = 0; l < csz; ++l) cval[cv0 + j * csz + l] = D[(int64_t)l * range + col];
  }
}

// per-block partial checksums of a cluster matrix: {sum v, sum |v|, sum colids, nonzeros}
template <typename V>
__global__ __launch_bounds__(kBlock) void k_cluster_checksum(const V *__restrict__ val, int64_t nnzv,
                                                             const int32_t *__restrict__ colids, int64_t nnzc,
                                                             double *__restrict__ part) {
  __shared__ double red[4][kBlock];
  double sv = 0, sa = 0, sc = 0, nzv = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nnzv; i += stride) {
    const double v = (double)val[i];
    sv += v;
    sa += fabs(v);
    nzv += v != 0.0 ? 1.0 : 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nnzc; i += stride) sc += (double)colids[i];
  red[0][threadIdx.x] = sv;
  red[1][threadIdx.x] = sa;
  red[2][threadIdx.x] = sc;
  red[3][threadIdx.x] = nzv;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

void spgemm_plan_create(const spops_cluster_s &A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *brp64,
                        const int64_t *bcol64, const void *bval, spops_spgemm_plan_s &P, hipStream_t s) {
  const int vt = A.value_type;
  const int64_t nc = A.rows;
  const size_t vb = value_bytes(vt);
  Timer tm(s);
  SPOPS_REQUIRE(b_cols < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT, "B must have < 2^31 columns");
  P.A = &A;
  P.b_rows = b_rows;
  P.b_cols = b_cols;
  P.b_nnz = b_nnz;
  P.brp64 = brp64;
  P.bval = bval;
  // B with int32 columns, rows sorted (the large-cluster kernels walk B rows in column order)
  DevBuf flag;
  P.bcol32.alloc((size_t)std::max<int64_t>(b_nnz, 1) * 4);
  flag.alloc(16);
  SPOPS_HIP_CHECK(hipMemsetAsync(flag.p, 0, 16, s));
  if (b_nnz > 0) {
    hipLaunchKernelGGL(k_narrow, dim3(nblk(b_nnz, kBlock)), dim3(kBlock), 0, s, bcol64, b_nnz, P.bcol32.as<int32_t>());
    hipLaunchKernelGGL(k_rows_unsorted, dim3(nblk(b_rows, kBlock / kWave)), dim3(kBlock), 0, s, brp64, b_rows,
                       P.bcol32.as<int32_t>(), flag.as<int>());
  }
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipMemcpyAsync(&P.unsorted, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  if (P.unsorted) {
    SPOPS_REQUIRE(b_rows < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "B must have < 2^31 rows");
    DevBuf keys, idx;
    keys.alloc((size_t)b_nnz * 8);
    idx.alloc((size_t)b_nnz * 8);
    hipLaunchKernelGGL(k_row_keys, dim3(nblk(b_rows, kBlock / kWave)), dim3(kBlock), 0, s, brp64, b_rows,
                       P.bcol32.as<int32_t>(), keys.as<uint64_t>(), idx.as<uint64_t>());
    SPOPS_HIP_CHECK(hipGetLastError());
    radix_sort_u64(keys.as<uint64_t>(), idx.p, 8, b_nnz, 0, 64, s);
    P.bsorted_val.alloc((size_t)b_nnz * vb);
    if (vt == SPOPS_VALUE_F64)
      hipLaunchKernelGGL(k_row_sorted_fill<double>, dim3(nblk(b_nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(),
                         idx.as<uint64_t>(), b_nnz, static_cast<const double *>(bval), P.bcol32.as<int32_t>(),
                         P.bsorted_val.as<double>());
    else
      hipLaunchKernelGGL(k_row_sorted_fill<float>, dim3(nblk(b_nnz, kBlock)), dim3(kBlock), 0, s, keys.as<uint64_t>(),
                         idx.as<uint64_t>(), b_nnz, static_cast<const float *>(bval), P.bcol32.as<int32_t>(),
                         P.bsorted_val.as<float>());
    SPOPS_HIP_CHECK(hipGetLastError());
    P.bval = P.bsorted_val.p;
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  }
  // work estimate and bins
  DevBuf flop;
  flop.alloc((size_t)std::max<int64_t>(nc, 1) * 8);
  if (nc > 0)
    hipLaunchKernelGGL(k_cluster_flop, dim3(nblk(nc, kBlock / kWave)), dim3(kBlock), 0, s, A.rowptr.as<int64_t>(),
                       A.colids.as<int32_t>(), brp64, nc, flop.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  P.hflop.assign(nc, 0);
  P.hcsz.assign(nc, 0);
  if (nc > 0) {
    SPOPS_HIP_CHECK(hipMemcpyAsync(P.hflop.data(), flop.p, (size_t)nc * 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipMemcpyAsync(P.hcsz.data(), A.cluster_sz.p, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
  }
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  P.small.clear();
  P.large.clear();
  P.products = 0;
  for (int64_t c = 0; c < nc; ++c) {
    P.products += P.hflop[c];
    if (P.hflop[c] == 0) continue;
    if (P.hflop[c] <= kSmallFlop && P.hflop[c] * P.hcsz[c] <= kSmallAcc) P.small.push_back((int32_t)c);
    else P.large.push_back((int32_t)c);
  }
  // large clusters in descending work, so that the longest start first (ties in ascending id)
  sort_by_work_desc(P.large, P.hflop);
  P.dsmall.upload(P.small);
  P.harp.assign(nc + 1, 0);
  if (nc > 0)
    SPOPS_HIP_CHECK(hipMemcpyAsync(P.harp.data(), A.rowptr.p, (size_t)(nc + 1) * 8, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  // heavy clusters (by decreasing work, while the part slices and kept bitmaps fit their
  // budgets): parts of at most kTileCursors A columns and about kHeavyPartProducts products
  const int64_t words = (b_cols + 31) / 32;
  std::vector<int32_t> light;
  std::vector<int4> heavy;                 // {cluster, first slice, end slice, bitmap}
  std::vector<HeavyPart> parts;
  int64_t nslices = 0;
  P.heavy_id.assign(nc, -1);
  for (int32_t c : P.large) {
    const int64_t na = P.harp[c + 1] - P.harp[c];
    int64_t np = std::max<int64_t>((P.hflop[c] + kHeavyPartProducts - 1) / kHeavyPartProducts,
                                   (na + kTileCursors - 1) / kTileCursors);
    np = std::min<int64_t>(np, na);
    const bool is_heavy = P.hflop[c] >= g_heavy_products && na > 0 && words > 0 &&
                          (nslices + np) * words * 4 <= kHeavySliceBytes &&
                          ((int64_t)heavy.size() + 1) * words * 4 <= kHeavyBitmapBytes;
    if (!is_heavy) {
      light.push_back(c);
      continue;
    }
    P.heavy_id[c] = (int32_t)heavy.size();
    heavy.push_back(make_int4(c, (int)nslices, (int)(nslices + np), (int)heavy.size()));
    for (int64_t i = 0; i < np; ++i)
      parts.push_back(HeavyPart{P.harp[c] + na * i / np, P.harp[c] + na * (i + 1) / np, (int32_t)(nslices + i), 0});
    nslices += np;
  }
  std::vector<int2> heavy_list;
  for (const int4 &h : heavy) heavy_list.push_back(make_int2(h.x, h.w));
  P.dlight.upload(light);
  P.dheavy.upload(heavy_list);
  P.nlight = (int64_t)light.size();
  P.nheavy = (int64_t)heavy.size();
  P.heavy_words = words;
  const size_t nb = (size_t)std::max<int64_t>(nc, 1) * 8;
  P.nz.alloc(nb);
  P.cmin.alloc(nb);
  P.cmax.alloc(nb);
  SPOPS_HIP_CHECK(hipMemsetAsync(P.nz.p, 0, nb, s));
  SPOPS_HIP_CHECK(hipMemsetAsync(P.cmin.p, 0, nb, s));
  SPOPS_HIP_CHECK(hipMemsetAsync(P.cmax.p, 0, nb, s));
  P.ms_setup = tm.stop();
  // ---- symbolic, every cluster
  const int b32 = (int)(b_nnz < (int64_t)INT32_MAX);
  // B-row cursors (position, end) per cluster column, for clusters of more than kTileCursors
  // columns (the bitmap passes keep the others in LDS)
  bool wide_cluster = false;
  for (int32_t c : light) wide_cluster |= P.harp[c + 1] - P.harp[c] > kTileCursors;
  if (b32 && b_cols > kLargeTileBits && wide_cluster) P.gcur.alloc((size_t)A.nnzc * 8);
  if (!P.small.empty())
    hipLaunchKernelGGL(k_sym_small, dim3((unsigned)P.small.size()), dim3(kBlock), 0, s, P.dsmall.as<int32_t>(),
                       (int64_t)P.small.size(), A.rowptr.as<int64_t>(), A.colids.as<int32_t>(), brp64,
                       P.bcol32.as<int32_t>(), P.nz.as<int64_t>(), P.cmin.as<int64_t>(), P.cmax.as<int64_t>());
  DevBuf dparts, dheavy4, slices;
  if (!heavy.empty()) {
    dparts.upload(parts);
    dheavy4.upload(heavy);
    slices.alloc((size_t)nslices * words * 4);
    P.heavy_bm.alloc((size_t)heavy.size() * words * 4);
    hipLaunchKernelGGL(k_heavy_part, dim3((unsigned)parts.size()), dim3(kBitmapBlock), 0, s, dparts.as<HeavyPart>(),
                       A.colids.as<int32_t>(), brp64, P.bcol32.as<int32_t>(), (int32_t)b_cols, words,
                       slices.as<uint32_t>(), b32);
    hipLaunchKernelGGL(k_heavy_merge, dim3((unsigned)heavy.size()), dim3(kBlock), 0, s, dheavy4.as<int4>(),
                       slices.as<uint32_t>(), words, P.heavy_bm.as<uint32_t>(), P.nz.as<int64_t>(),
                       P.cmin.as<int64_t>(), P.cmax.as<int64_t>());
  }
  if (!light.empty())
    hipLaunchKernelGGL(k_sym_large, dim3((unsigned)light.size()), dim3(kBitmapBlock), 0, s, P.dlight.as<int32_t>(),
                       A.rowptr.as<int64_t>(), A.colids.as<int32_t>(), brp64, P.bcol32.as<int32_t>(), (int32_t)b_cols,
                       P.nz.as<int64_t>(), P.cmin.as<int64_t>(), P.cmax.as<int64_t>(), b32, P.gcur.as<int32_t>(),
                       P.gcur.p ? P.gcur.as<int32_t>() + A.nnzc : nullptr);
  SPOPS_HIP_CHECK(hipGetLastError());
  P.hnz.assign(nc, 0);
  if (nc > 0) SPOPS_HIP_CHECK(hipMemcpyAsync(P.hnz.data(), P.nz.p, (size_t)nc * 8, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  P.nnzc = P.nnzv = 0;
  for (int64_t c = 0; c < nc; ++c) {
    P.nnzc += P.hnz[c];
    P.nnzv += P.hnz[c] * P.hcsz[c];
  }
  // clusters with very many products accumulate densely (see k_dense_products)
  P.hcmin.assign(nc, 0);
  P.hcmax.assign(nc, 0);
  if (nc > 0) {
    SPOPS_HIP_CHECK(hipMemcpyAsync(P.hcmin.data(), P.cmin.p, (size_t)nc * 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipMemcpyAsync(P.hcmax.data(), P.cmax.p, (size_t)nc * 8, hipMemcpyDeviceToHost, s));
  }
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  P.dense.assign(nc, 0);
  for (int32_t c : P.large) {
    const int64_t bytes = (P.hcmax[c] - P.hcmin[c] + 1) * P.hcsz[c] * (int64_t)vb;
    P.dense[c] = P.hflop[c] >= g_dense_products && P.hnz[c] > 0 && bytes <= kDenseScratch;
  }
  P.ms_symbolic = tm.stop();
}

namespace {
void ensure(DevBuf &b, size_t bytes) {
  if (b.bytes < bytes || b.p == nullptr) b.alloc(bytes);
}
}  // namespace

void spgemm_plan_numeric(spops_spgemm_plan_s &P, int64_t c0, int64_t c1, spops_cluster_s &C,
                         int64_t stats[kSpgemmStats], hipStream_t s) {
  const spops_cluster_s &A = *P.A;
  const int vt = A.value_type;
  const int64_t nc = A.rows;
  SPOPS_REQUIRE(0 <= c0 && c0 <= c1 && c1 <= nc, SPOPS_ERROR_INVALID_VALUE, "cluster range [%lld, %lld) outside [0, %lld)",
                (long long)c0, (long long)c1, (long long)nc);
  const int64_t np = c1 - c0;
  const size_t vb = value_bytes(vt);
  const bool whole = c0 == 0 && c1 == nc;
  Timer tm(s);
  // the panel's clusters with work, in the plan's bin order
  std::vector<int32_t> psmall, plarge, pregular, pdense;
  DevBuf dsmall;
  if (!whole) {
    for (int32_t c : P.small) if (c >= c0 && c < c1) psmall.push_back(c);
    for (int32_t c : P.large) if (c >= c0 && c < c1) plarge.push_back(c);
    dsmall.upload(psmall);
  }
  const std::vector<int32_t> &small = whole ? P.small : psmall;
  const std::vector<int32_t> &large = whole ? P.large : plarge;
  const int32_t *lsmall = whole ? P.dsmall.as<int32_t>() : dsmall.as<int32_t>();
  // the emit pass: large clusters without a kept bitmap, and the heavy ones
  std::vector<int32_t> plight;
  std::vector<int2> pheavy;
  DevBuf dlight, dheavy;
  if (!whole) {
    for (int32_t c : plarge) {
      if (P.heavy_id[c] >= 0) pheavy.push_back(make_int2(c, P.heavy_id[c]));
      else plight.push_back(c);
    }
    dlight.upload(plight);
    dheavy.upload(pheavy);
  }
  const int64_t nlight = whole ? P.nlight : (int64_t)plight.size();
  const int64_t nheavy = whole ? P.nheavy : (int64_t)pheavy.size();
  const int32_t *llight = whole ? P.dlight.as<int32_t>() : dlight.as<int32_t>();
  const int2 *lheavy = whole ? P.dheavy.as<int2>() : dheavy.as<int2>();
  for (int32_t c : large) (P.dense[c] ? pdense : pregular).push_back(c);
  const bool split = !pdense.empty();
  const std::vector<int32_t> &regular = split ? pregular : large;
  // tasks of the large-cluster numeric kernels: {cluster, first chunk, end chunk, 0}, about
  // task_products products each (clusters in decreasing work order)
  // (clusters with more than kLargeCursors columns and B offsets of 32 bits: the variant with
  // kLargeCursorsWide cursors)
  std::vector<int4> regular4, regular16, wide4, wide16;
  const bool b32 = P.b_nnz < (int64_t)INT32_MAX;
  auto pieces_of = [&](int32_t c, int64_t per_task, int64_t *chunks_out) {
    const int cr = P.hcsz[c] <= 4 ? 4 : kMaxClusterRows;
    const int64_t chunks = (P.hnz[c] + large_chunk_cols(P.hcsz[c], cr) - 1) / large_chunk_cols(P.hcsz[c], cr);
    if (chunks_out) *chunks_out = chunks;
    return std::max<int64_t>(1, std::min<int64_t>(chunks, P.hflop[c] / per_task));
  };
  // products per task: the tuned value, else kLargeTaskProducts halved (down to
  // kLargeTaskProductsMin) while the call has fewer than kLargeTasksMin tasks -- the smaller
  // the product, the more the tail of the launch matters
  int64_t task_products = g_task_products;
  if (task_products <= 0) {
    constexpr int kSizes = 8;                  // candidate sizes kLargeTaskProducts >> i
    int64_t tasks[kSizes] = {0};
    int nsz = 1;
    while (nsz < kSizes && (kLargeTaskProducts >> nsz) >= kLargeTaskProductsMin) ++nsz;
    for (int32_t c : regular)
      for (int i = 0; i < nsz; ++i) tasks[i] += pieces_of(c, kLargeTaskProducts >> i, nullptr);
    int i = 0;
    while (i + 1 < nsz && tasks[i] < kLargeTasksMin) ++i;
    task_products = kLargeTaskProducts >> i;
  }
  // clusters of kLargeCursors+1..kLargeCursorsWide A columns in the variant with
  // kLargeCursorsWide cursors (instead of two binary searches per row and chunk).  It runs at
  // a quarter of the regular kernel's waves, so per call (mode 2) it takes them when they are
  // at least kWideMinTasks tasks, 16 per workgroup slot of the variant (measured: R-MAT 20,
  // 1.8-39 K such tasks per panel, +4.6 % with the variant on all panels; R-MAT 16 and 18,
  // 3.5 K and 5.3 K tasks: -4 % and 0 %)
  auto is_wide = [&](int32_t c) {
    const int64_t na = P.harp[c + 1] - P.harp[c];
    return b32 && na > kLargeCursors && na <= kLargeCursorsWide;
  };
  bool use_wide = g_wide_cursors == 1;
  if (g_wide_cursors == 2) {
    int64_t wide_tasks = 0;
    for (int32_t c : regular)
      if (is_wide(c)) wide_tasks += pieces_of(c, task_products, nullptr);
    use_wide = wide_tasks >= kWideMinTasks;
  }
  for (int32_t c : regular) {
    const int cr = P.hcsz[c] <= 4 ? 4 : kMaxClusterRows;
    int64_t chunks = 0;
    const int64_t pieces = pieces_of(c, task_products, &chunks);
    const bool wide = use_wide && is_wide(c);
    std::vector<int4> &to = wide ? (cr == 4 ? wide4 : wide16) : (cr == 4 ? regular4 : regular16);
    for (int64_t i = 0; i < pieces; ++i)
      to.push_back(make_int4(c, (int)(chunks * i / pieces), (int)(chunks * (i + 1) / pieces), 0));
  }
  DevBuf dregular4, dregular16, dwide4, dwide16;
  dregular4.upload(regular4);
  dregular16.upload(regular16);
  dwide4.upload(wide4);
  dwide16.upload(wide16);
  int64_t nnzc = 0, nnzv = 0, csr_rows = 0;
  for (int64_t c = c0; c < c1; ++c) {
    nnzc += P.hnz[c];
    nnzv += P.hnz[c] * P.hcsz[c];
    csr_rows += P.hcsz[c];
  }
  C.csr_rows = whole ? A.csr_rows : csr_rows;
  C.rows = np;
  C.cols = P.b_cols;
  C.value_type = vt;
  C.max_cluster_sz = A.max_cluster_sz;
  C.nnzc = nnzc;
  C.nnzv = nnzv;
  ensure(C.rowptr, (size_t)(np + 1) * 8);
  ensure(C.rowptr_val, (size_t)(np + 1) * 8);
  {
    DevBuf nzp;
    nzp.alloc((size_t)(np + 1) * 8);
    SPOPS_HIP_CHECK(hipMemsetAsync(nzp.p, 0, (size_t)(np + 1) * 8, s));
    if (np > 0)
      SPOPS_HIP_CHECK(hipMemcpyAsync(nzp.p, P.nz.as<int64_t>() + c0, (size_t)np * 8, hipMemcpyDeviceToDevice, s));
    int64_t t = 0;
    exclusive_scan_i64(nzp.as<int64_t>(), C.rowptr.as<int64_t>(), np + 1, s, &t);
    if (np > 0)
      hipLaunchKernelGGL(k_mul_csz, dim3(nblk(np, kBlock)), dim3(kBlock), 0, s, P.nz.as<int64_t>() + c0,
                         A.cluster_sz.as<int32_t>() + c0, np, nzp.as<int64_t>());
    exclusive_scan_i64(nzp.as<int64_t>(), C.rowptr_val.as<int64_t>(), np + 1, s, &t);
  }
  // per-cluster symbolic results and sizes of the panel
  ensure(C.cluster_sz, (size_t)std::max<int64_t>(np, 1) * 4);
  ensure(C.nzc, (size_t)std::max<int64_t>(np, 1) * 8);
  ensure(C.cmin, (size_t)std::max<int64_t>(np, 1) * 8);
  ensure(C.cmax, (size_t)std::max<int64_t>(np, 1) * 8);
  if (np > 0) {
    SPOPS_HIP_CHECK(hipMemcpyAsync(C.cluster_sz.p, A.cluster_sz.as<int32_t>() + c0, (size_t)np * 4,
                                   hipMemcpyDeviceToDevice, s));
    SPOPS_HIP_CHECK(hipMemcpyAsync(C.nzc.p, P.nz.as<int64_t>() + c0, (size_t)np * 8, hipMemcpyDeviceToDevice, s));
    SPOPS_HIP_CHECK(hipMemcpyAsync(C.cmin.p, P.cmin.as<int64_t>() + c0, (size_t)np * 8, hipMemcpyDeviceToDevice, s));
    SPOPS_HIP_CHECK(hipMemcpyAsync(C.cmax.p, P.cmax.as<int64_t>() + c0, (size_t)np * 8, hipMemcpyDeviceToDevice, s));
  }
  double ms_scan = tm.stop();
  // ---- allocate output (the reference's "malloc" phase); every entry is written below
  ensure(C.colids, (size_t)std::max<int64_t>(nnzc, 1) * 4);
  ensure(C.values, (size_t)std::max<int64_t>(nnzv, 1) * vb);
  double ms_malloc = tm.stop();
  // ---- numeric.  The kernels index the output's cluster pointers by the global cluster id c,
  // so they get the panel's arrays shifted by c0 (only ids in [c0, c1) are touched).
  int64_t *crp = C.rowptr.as<int64_t>() - c0;
  int64_t *crpv = C.rowptr_val.as<int64_t>() - c0;
  if (nlight > 0)
    hipLaunchKernelGGL(k_emit_large, dim3((unsigned)nlight), dim3(kBitmapBlock), 0, s, llight, A.rowptr.as<int64_t>(),
                       A.colids.as<int32_t>(), P.brp64, P.bcol32.as<int32_t>(), crp, P.cmin.as<int64_t>(),
                       P.cmax.as<int64_t>(), C.colids.as<int32_t>(), (int)(P.b_nnz < (int64_t)INT32_MAX),
                       P.gcur.as<int32_t>(), P.gcur.p ? P.gcur.as<int32_t>() + A.nnzc : nullptr);
  if (nheavy > 0)
    hipLaunchKernelGGL(k_heavy_emit, dim3((unsigned)nheavy), dim3(kBlock), 0, s, lheavy, P.heavy_bm.as<uint32_t>(),
                       P.heavy_words, crp, P.cmin.as<int64_t>(), P.cmax.as<int64_t>(), C.colids.as<int32_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  auto numeric = [&](auto tag) {
    using V = decltype(tag);
    if (!small.empty())
      hipLaunchKernelGGL(k_num_small<V>, dim3((unsigned)small.size()), dim3(kBlock), 0, s, lsmall,
                         (int64_t)small.size(), A.rowptr.as<int64_t>(), A.rowptr_val.as<int64_t>(),
                         A.cluster_sz.as<int32_t>(), A.colids.as<int32_t>(), A.values.as<V>(), P.brp64,
                         P.bcol32.as<int32_t>(), static_cast<const V *>(P.bval), crp, crpv, C.colids.as<int32_t>(),
                         C.values.as<V>());
    // clusters of at most 4 rows and the others: kernels holding 4 / kMaxClusterRows A values
    auto large_kernel = [&](auto kern, const std::vector<int4> &ids, const int4 *dids) {
      if (!ids.empty())
        hipLaunchKernelGGL(kern, dim3((unsigned)ids.size()), dim3(kLargeBlock), 0, s, dids,
                           A.rowptr.as<int64_t>(), A.rowptr_val.as<int64_t>(), A.cluster_sz.as<int32_t>(),
                           A.colids.as<int32_t>(), A.values.as<V>(), P.brp64, P.bcol32.as<int32_t>(),
                           static_cast<const V *>(P.bval), crp, crpv, C.colids.as<int32_t>(), C.values.as<V>(),
                           P.b_nnz < (int64_t)INT32_MAX);
    };
    large_kernel(k_num_large<V, 4, kLargeCursorsWide>, wide4, dwide4.as<int4>());
    large_kernel(k_num_large<V, kMaxClusterRows, kLargeCursorsWide>, wide16, dwide16.as<int4>());
    large_kernel(k_num_large<V, 4, kLargeCursors>, regular4, dregular4.as<int4>());
    large_kernel(k_num_large<V, kMaxClusterRows, kLargeCursors>, regular16, dregular16.as<int4>());
    // dense clusters in batches whose accumulators fit kDenseScratch
    size_t i = 0;
    while (i < pdense.size()) {
      std::vector<DenseTask> tasks;
      std::vector<int32_t> blist;
      std::vector<int64_t> boff;
      int64_t used = 0, maxnz = 0;
      while (i < pdense.size()) {
        const int32_t c = pdense[i];
        const int64_t n = (P.hcmax[c] - P.hcmin[c] + 1) * P.hcsz[c];
        if (!blist.empty() && (used + n) * (int64_t)sizeof(V) > kDenseScratch) break;
        blist.push_back(c);
        boff.push_back(used);
        for (int64_t t = P.harp[c]; t < P.harp[c + 1]; t += kDenseTaskCols)
          tasks.push_back(DenseTask{c, t, std::min<int64_t>(t + kDenseTaskCols, P.harp[c + 1]), used});
        used += n;
        maxnz = std::max(maxnz, P.hnz[c]);
        ++i;
      }
      ensure(P.scratch, (size_t)std::max<int64_t>(used, 1) * sizeof(V));
      DevBuf dtasks, dlist, doff;
      dtasks.upload(tasks);
      dlist.upload(blist);
      doff.upload(boff);
      SPOPS_HIP_CHECK(hipMemsetAsync(P.scratch.p, 0, (size_t)used * sizeof(V), s));
      if (!tasks.empty())
        hipLaunchKernelGGL(k_dense_products<V>, dim3((unsigned)tasks.size()), dim3(kBlock), 0, s,
                           dtasks.as<DenseTask>(), A.rowptr.as<int64_t>(), A.rowptr_val.as<int64_t>(),
                           A.cluster_sz.as<int32_t>(), A.colids.as<int32_t>(), A.values.as<V>(), P.brp64,
                           P.bcol32.as<int32_t>(), static_cast<const V *>(P.bval), P.cmin.as<int64_t>(),
                           P.cmax.as<int64_t>(), P.scratch.as<V>());
      SPOPS_HIP_CHECK(hipGetLastError());
      const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>((maxnz + kBlock - 1) / kBlock, 1), 1024);
      hipLaunchKernelGGL(k_dense_gather<V>, dim3(gx, (unsigned)blist.size()), dim3(kBlock), 0, s, dlist.as<int32_t>(),
                         doff.as<int64_t>(), A.cluster_sz.as<int32_t>(), P.cmin.as<int64_t>(), P.cmax.as<int64_t>(),
                         P.scratch.as<V>(), crp, crpv, C.colids.as<int32_t>(), C.values.as<V>());
      SPOPS_HIP_CHECK(hipGetLastError());
      SPOPS_HIP_CHECK(hipStreamSynchronize(s));    // the batch's task lists are freed next
    }
  };
  if (vt == SPOPS_VALUE_F64) numeric(double(0));
  else numeric(float(0));
  SPOPS_HIP_CHECK(hipGetLastError());
  double ms_numeric = tm.stop();
  int64_t products = 0;
  for (int64_t c = c0; c < c1; ++c) products += P.hflop[c];
  stats[0] = P.unsorted;
  stats[1] = nnzc;
  stats[2] = nnzv;
  stats[3] = (int64_t)small.size();
  stats[4] = (int64_t)large.size();
  stats[10] = (int64_t)pdense.size();
  stats[11] = (int64_t)(regular4.size() + regular16.size());
  stats[12] = (int64_t)(wide4.size() + wide16.size());
  stats[5] = (int64_t)((whole ? P.ms_setup : 0.0) * 1e3);
  stats[6] = (int64_t)((whole ? P.ms_symbolic : 0.0) * 1e3 + ms_scan * 1e3);
  stats[7] = (int64_t)(ms_malloc * 1e3);
  stats[8] = (int64_t)(ms_numeric * 1e3);
  stats[9] = products;
}

void lespgemm(const spops_cluster_s &A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *brp64,
              const int64_t *bcol64, const void *bval, spops_cluster_s &C, int64_t stats[kSpgemmStats], hipStream_t s) {
  spops_spgemm_plan_s P;
  spgemm_plan_create(A, b_rows, b_cols, b_nnz, brp64, bcol64, bval, P, s);
  spgemm_plan_numeric(P, 0, A.rows, C, stats, s);
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));    // the plan's buffers are freed on return
}

void cluster_to_csr(const spops_cluster_s &C, const spops_cluster_s &A, int64_t &nnz, DevBuf &rowptr, DevBuf &col,
                    DevBuf &val, hipStream_t s) {
  const int64_t n = C.csr_rows;
  DevBuf cnt;
  cnt.alloc((size_t)(n + 1) * 8);
  SPOPS_HIP_CHECK(hipMemsetAsync(cnt.p, 0, (size_t)(n + 1) * 8, s));
  const unsigned grid = nblk(std::max<int64_t>(n, 1), kBlock / kWave);
  if (C.value_type == SPOPS_VALUE_F64)
    hipLaunchKernelGGL(k_c2csr_count<double>, dim3(grid), dim3(kBlock), 0, s, A.offsets.as<int64_t>(), A.perm.as<int64_t>(),
                       A.cid.as<int32_t>(), A.lrow.as<int32_t>(), C.cluster_sz.as<int32_t>(), C.rowptr.as<int64_t>(),
                       C.rowptr_val.as<int64_t>(), C.values.as<double>(), n, cnt.as<int64_t>());
  else
    hipLaunchKernelGGL(k_c2csr_count<float>, dim3(grid), dim3(kBlock), 0, s, A.offsets.as<int64_t>(), A.perm.as<int64_t>(),
                       A.cid.as<int32_t>(), A.lrow.as<int32_t>(), C.cluster_sz.as<int32_t>(), C.rowptr.as<int64_t>(),
                       C.rowptr_val.as<int64_t>(), C.values.as<float>(), n, cnt.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  rowptr.alloc((size_t)(n + 1) * 8);
  exclusive_scan_i64(cnt.as<int64_t>(), rowptr.as<int64_t>(), n + 1, s, &nnz);
  col.alloc((size_t)std::max<int64_t>(nnz, 1) * 8);
  val.alloc((size_t)std::max<int64_t>(nnz, 1) * value_bytes(C.value_type));
  if (C.value_type == SPOPS_VALUE_F64)
    hipLaunchKernelGGL(k_c2csr_fill<double>, dim3(grid), dim3(kBlock), 0, s, A.perm.as<int64_t>(), A.cid.as<int32_t>(),
                       A.lrow.as<int32_t>(), C.cluster_sz.as<int32_t>(), C.rowptr.as<int64_t>(), C.rowptr_val.as<int64_t>(),
                       C.colids.as<int32_t>(), C.values.as<double>(), n, rowptr.as<int64_t>(), col.as<int64_t>(),
                       val.as<double>());
  else
    hipLaunchKernelGGL(k_c2csr_fill<float>, dim3(grid), dim3(kBlock), 0, s, A.perm.as<int64_t>(), A.cid.as<int32_t>(),
                       A.lrow.as<int32_t>(), C.cluster_sz.as<int32_t>(), C.rowptr.as<int64_t>(), C.rowptr_val.as<int64_t>(),
                       C.colids.as<int32_t>(), C.values.as<float>(), n, rowptr.as<int64_t>(), col.as<int64_t>(),
                       val.as<float>());
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace spops

// =============================================================================================
// C ABI
// =============================================================================================
using namespace spops;

namespace {
unsigned nblocks(int64_t n, int per) { return grid_blocks(n, per); }
}  // namespace

extern "C" int spops_cluster_create(spops_cluster *out, int64_t n_rows, int64_t n_cols, int64_t nnz,
                                    const int64_t *d_rowptr, const int64_t *d_col, const void *d_values, int value_type,
                                    const int64_t *perm, int64_t nclusters, const int64_t *offsets, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    SPOPS_REQUIRE(value_type == SPOPS_VALUE_F64 || value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "cluster values must be F64 or F32");
    SPOPS_REQUIRE(n_rows >= 0 && nnz >= 0 && nclusters >= 0, SPOPS_ERROR_INVALID_VALUE, "negative size");
    SPOPS_REQUIRE(n_rows < ((int64_t)1 << 31) && n_cols < ((int64_t)1 << 31) && nclusters < ((int64_t)1 << 31),
                  SPOPS_ERROR_LIMIT, "rows, columns and clusters must be < 2^31");
    SPOPS_REQUIRE(nnz < ((int64_t)1 << 40), SPOPS_ERROR_LIMIT, "nnz must be < 2^40");
    SPOPS_REQUIRE(perm && offsets && (n_rows == 0 || (d_rowptr && d_col && (nnz == 0 || d_values))),
                  SPOPS_ERROR_INVALID_VALUE, "null pointer");
    // csr_to_vlength_cluster's checks (include/sparse_conversion.h:1656-1667)
    SPOPS_REQUIRE(nclusters >= 1 || n_rows == 0, SPOPS_ERROR_INVALID_VALUE, "offsets must have at least 2 elements");
    SPOPS_REQUIRE(offsets[0] == 0 && offsets[nclusters] == n_rows, SPOPS_ERROR_INVALID_VALUE,
                  "Invalid offsets: must be [0, ..., csr.num_rows]");
    for (int64_t c = 0; c < nclusters; ++c)
      SPOPS_REQUIRE(offsets[c + 1] > offsets[c], SPOPS_ERROR_INVALID_VALUE, "empty or decreasing cluster %lld",
                    (long long)c);
    std::vector<char> seen(n_rows, 0);
    for (int64_t t = 0; t < n_rows; ++t) {
      SPOPS_REQUIRE(perm[t] >= 0 && perm[t] < n_rows && !seen[perm[t]], SPOPS_ERROR_INVALID_VALUE,
                    "perm is not a permutation (entry %lld)", (long long)t);
      seen[perm[t]] = 1;
    }
    auto A = std::make_unique<spops_cluster_s>();
    cluster_build(*A, n_rows, n_cols, nnz, d_rowptr, d_col, d_values, value_type, perm, nclusters, offsets,
                  static_cast<hipStream_t>(stream));
    *out = A.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_info(spops_cluster A, int64_t info[8]) {
  return guarded([&]() {
    SPOPS_REQUIRE(A && info, SPOPS_ERROR_INVALID_VALUE, "null argument");
    info[0] = A->rows; info[1] = A->csr_rows; info[2] = A->cols; info[3] = A->nnzc; info[4] = A->nnzv;
    info[5] = A->max_cluster_sz; info[6] = A->value_type;
    info[7] = (int64_t)(A->cluster_sz.bytes + A->rowptr.bytes + A->rowptr_val.bytes + A->colids.bytes + A->values.bytes);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_download(spops_cluster A, int64_t *cluster_sz, int64_t *rowptr, int64_t *rowptr_val,
                                      int64_t *colids, void *values) {
  return guarded([&]() {
    SPOPS_REQUIRE(A, SPOPS_ERROR_INVALID_VALUE, "null cluster matrix");
    const size_t nc = (size_t)A->rows;
    if (cluster_sz && nc) {
      std::vector<int32_t> t(nc);
      SPOPS_HIP_CHECK(hipMemcpy(t.data(), A->cluster_sz.p, nc * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nc; ++i) cluster_sz[i] = t[i];
    }
    if (rowptr) SPOPS_HIP_CHECK(hipMemcpy(rowptr, A->rowptr.p, (nc + 1) * 8, hipMemcpyDeviceToHost));
    if (rowptr_val) SPOPS_HIP_CHECK(hipMemcpy(rowptr_val, A->rowptr_val.p, (nc + 1) * 8, hipMemcpyDeviceToHost));
    if (colids && A->nnzc) {
      std::vector<int32_t> t((size_t)A->nnzc);
      SPOPS_HIP_CHECK(hipMemcpy(t.data(), A->colids.p, t.size() * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < t.size(); ++i) colids[i] = t[i];
    }
    if (values && A->nnzv)
      SPOPS_HIP_CHECK(hipMemcpy(values, A->values.p, (size_t)A->nnzv * value_bytes(A->value_type), hipMemcpyDeviceToHost));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_symbolic_download(spops_cluster C, int64_t *nz, int64_t *cmin, int64_t *cmax) {
  return guarded([&]() {
    SPOPS_REQUIRE(C && C->nzc.p, SPOPS_ERROR_INVALID_VALUE, "not a product cluster matrix");
    const size_t nc = (size_t)C->rows;
    if (!nc) return SPOPS_SUCCESS;
    if (nz) SPOPS_HIP_CHECK(hipMemcpy(nz, C->nzc.p, nc * 8, hipMemcpyDeviceToHost));
    if (cmin) SPOPS_HIP_CHECK(hipMemcpy(cmin, C->cmin.p, nc * 8, hipMemcpyDeviceToHost));
    if (cmax) SPOPS_HIP_CHECK(hipMemcpy(cmax, C->cmax.p, nc * 8, hipMemcpyDeviceToHost));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_destroy(spops_cluster A) {
  delete A;
  return SPOPS_SUCCESS;
}

extern "C" int spops_lespgemm(spops_cluster A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *d_browptr,
                              const int64_t *d_bcol, const void *d_bval, spops_cluster *C_out, int64_t stats[16],
                              void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(A && C_out, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(A->cols == b_rows, SPOPS_ERROR_INVALID_VALUE, "A.cols (%lld) != B.rows (%lld)", (long long)A->cols,
                  (long long)b_rows);
    SPOPS_REQUIRE(b_cols < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "B columns must be < 2^31");
    SPOPS_REQUIRE(b_rows == 0 || (d_browptr && (b_nnz == 0 || (d_bcol && d_bval))), SPOPS_ERROR_INVALID_VALUE,
                  "null B array");
    int64_t st[kSpgemmStats] = {0};
    auto C = std::make_unique<spops_cluster_s>();
    lespgemm(*A, b_rows, b_cols, b_nnz, d_browptr, d_bcol, d_bval, *C, st, static_cast<hipStream_t>(stream));
    if (stats)
      for (int i = 0; i < 16; ++i) stats[i] = i < kSpgemmStats ? st[i] : 0;
    *C_out = C.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lespgemm_plan_create(spops_cluster A, int64_t b_rows, int64_t b_cols, int64_t b_nnz,
                                          const int64_t *d_browptr, const int64_t *d_bcol, const void *d_bval,
                                          spops_spgemm_plan *out, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(A && out, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(A->cols == b_rows, SPOPS_ERROR_INVALID_VALUE, "A.cols (%lld) != B.rows (%lld)", (long long)A->cols,
                  (long long)b_rows);
    SPOPS_REQUIRE(b_cols < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "B columns must be < 2^31");
    SPOPS_REQUIRE(b_rows == 0 || (d_browptr && (b_nnz == 0 || (d_bcol && d_bval))), SPOPS_ERROR_INVALID_VALUE,
                  "null B array");
    auto P = std::make_unique<spops_spgemm_plan_s>();
    spgemm_plan_create(*A, b_rows, b_cols, b_nnz, d_browptr, d_bcol, d_bval, *P, static_cast<hipStream_t>(stream));
    *out = P.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lespgemm_plan_info(spops_spgemm_plan P, int64_t info[8]) {
  return guarded([&]() {
    SPOPS_REQUIRE(P && info, SPOPS_ERROR_INVALID_VALUE, "null argument");
    info[0] = P->A->rows;
    info[1] = P->nnzc;
    info[2] = P->nnzv;
    info[3] = P->products;
    info[4] = (int64_t)P->small.size();
    info[5] = (int64_t)P->large.size();
    info[6] = (int64_t)(P->ms_setup * 1e3);
    info[7] = (int64_t)(P->ms_symbolic * 1e3);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lespgemm_plan_panels(spops_spgemm_plan P, int64_t budget_bytes, int64_t *npanels,
                                          int64_t *bounds) {
  return guarded([&]() {
    SPOPS_REQUIRE(P && npanels && bounds, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(budget_bytes > 0, SPOPS_ERROR_INVALID_VALUE, "budget must be positive");
    const int64_t nc = P->A->rows, vb = (int64_t)value_bytes(P->A->value_type);
    int64_t n = 0, acc = 0;
    bounds[0] = 0;
    for (int64_t c = 0; c < nc; ++c) {
      const int64_t b = P->hnz[c] * (4 + P->hcsz[c] * vb) + 32;
      if (acc > 0 && acc + b > budget_bytes) {
        bounds[++n] = c;
        acc = 0;
      }
      acc += b;
    }
    bounds[++n] = nc;
    if (nc == 0) n = 1;
    *npanels = n;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lespgemm_plan_numeric(spops_spgemm_plan P, int64_t c0, int64_t c1, spops_cluster *C_inout,
                                           int64_t stats[16], void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(P && C_inout, SPOPS_ERROR_INVALID_VALUE, "null argument");
    std::unique_ptr<spops_cluster_s> fresh;
    spops_cluster_s *C = *C_inout;
    if (!C) {
      fresh = std::make_unique<spops_cluster_s>();
      C = fresh.get();
    }
    int64_t st[kSpgemmStats] = {0};
    spgemm_plan_numeric(*P, c0, c1, *C, st, static_cast<hipStream_t>(stream));
    if (stats)
      for (int i = 0; i < 16; ++i) stats[i] = i < kSpgemmStats ? st[i] : 0;
    if (fresh) *C_inout = fresh.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_checksum(spops_cluster C, double out[4], void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(C && out, SPOPS_ERROR_INVALID_VALUE, "null argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    constexpr int grid = 1024;
    DevBuf part;
    part.alloc((size_t)grid * 4 * 8);
    if (C->value_type == SPOPS_VALUE_F64)
      hipLaunchKernelGGL(k_cluster_checksum<double>, dim3(grid), dim3(kBlock), 0, s, C->values.as<double>(), C->nnzv,
                         C->colids.as<int32_t>(), C->nnzc, part.as<double>());
    else
      hipLaunchKernelGGL(k_cluster_checksum<float>, dim3(grid), dim3(kBlock), 0, s, C->values.as<float>(), C->nnzv,
                         C->colids.as<int32_t>(), C->nnzc, part.as<double>());
    SPOPS_HIP_CHECK(hipGetLastError());
    std::vector<double> h((size_t)grid * 4);
    SPOPS_HIP_CHECK(hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    for (int q = 0; q < 4; ++q) out[q] = 0;
    for (int g = 0; g < grid; ++g)
      for (int q = 0; q < 4; ++q) out[q] += h[(size_t)g * 4 + q];
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lespgemm_plan_destroy(spops_spgemm_plan P) {
  if (P) (void)hipDeviceSynchronize();
  delete P;
  return SPOPS_SUCCESS;
}

extern "C" int spops_cluster_to_csr(spops_cluster C, spops_cluster A, int64_t *nnz, int64_t **d_rowptr, int64_t **d_col,
                                    void **d_val, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(C && A && nnz && d_rowptr && d_col && d_val, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(C->rows == A->rows && C->csr_rows == A->csr_rows && A->perm.p, SPOPS_ERROR_INVALID_VALUE,
                  "C must be the product of the clustered matrix A");
    DevBuf rp, col, val;
    cluster_to_csr(*C, *A, *nnz, rp, col, val, static_cast<hipStream_t>(stream));
    *d_rowptr = static_cast<int64_t *>(rp.p); rp.p = nullptr;
    *d_col = static_cast<int64_t *>(col.p); col.p = nullptr;
    *d_val = val.p; val.p = nullptr;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spgemm_flop(int64_t a_nnz, const int64_t *d_acol, int64_t b_rows, const int64_t *d_browptr,
                                 int64_t *flops, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(flops, SPOPS_ERROR_INVALID_VALUE, "null output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf acc;
    acc.alloc(16);
    SPOPS_HIP_CHECK(hipMemsetAsync(acc.p, 0, 16, s));
    if (a_nnz > 0)
      hipLaunchKernelGGL(k_ref_flop, dim3(nblocks(a_nnz, kBlock)), dim3(kBlock), 0, s, d_acol, a_nnz, b_rows, d_browptr,
                         acc.as<unsigned long long>());
    SPOPS_HIP_CHECK(hipGetLastError());
    unsigned long long f = 0;
    SPOPS_HIP_CHECK(hipMemcpyAsync(&f, acc.p, 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    *flops = 2 * (int64_t)f;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_cluster_create_from_host(spops_cluster *out, int64_t csr_rows, int64_t cols, int64_t nclusters,
                                              const int64_t *cluster_sz, const int64_t *rowptr,
                                              const int64_t *rowptr_val, const int64_t *colids, const void *values,
                                              int value_type) {
  return guarded([&]() {
    SPOPS_REQUIRE(out && cluster_sz && rowptr && rowptr_val, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(value_type == SPOPS_VALUE_F64 || value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "cluster values must be F64 or F32");
    SPOPS_REQUIRE(csr_rows < ((int64_t)1 << 31) && cols < ((int64_t)1 << 31) && nclusters <= csr_rows,
                  SPOPS_ERROR_LIMIT, "rows and columns must be < 2^31");
    const int64_t nnzc = rowptr[nclusters], nnzv = rowptr_val[nclusters];
    std::vector<int64_t> offsets(nclusters + 1, 0);
    for (int64_t c = 0; c < nclusters; ++c) {
      SPOPS_REQUIRE(cluster_sz[c] > 0 && rowptr[c + 1] >= rowptr[c] &&
                        rowptr_val[c + 1] - rowptr_val[c] == (rowptr[c + 1] - rowptr[c]) * cluster_sz[c],
                    SPOPS_ERROR_INVALID_VALUE, "inconsistent cluster %lld", (long long)c);
      offsets[c + 1] = offsets[c] + cluster_sz[c];
    }
    SPOPS_REQUIRE(offsets[nclusters] == csr_rows, SPOPS_ERROR_INVALID_VALUE, "cluster sizes must sum to csr_rows");
    for (int64_t t = 0; t < nnzc; ++t)
      SPOPS_REQUIRE(colids[t] >= 0 && colids[t] < cols, SPOPS_ERROR_INVALID_VALUE, "column out of range");
    auto A = std::make_unique<spops_cluster_s>();
    A->csr_rows = csr_rows; A->rows = nclusters; A->cols = cols; A->nnzc = nnzc; A->nnzv = nnzv;
    A->value_type = value_type;
    std::vector<int32_t> csz(cluster_sz, cluster_sz + nclusters), cid(csr_rows), lrow(csr_rows);
    A->max_cluster_sz = 0;
    for (int64_t c = 0; c < nclusters; ++c) {
      A->max_cluster_sz = std::max<int64_t>(A->max_cluster_sz, csz[c]);
      for (int64_t t = offsets[c]; t < offsets[c + 1]; ++t) { cid[t] = (int32_t)c; lrow[t] = (int32_t)(t - offsets[c]); }
    }
    std::vector<int64_t> perm(csr_rows);
    std::iota(perm.begin(), perm.end(), 0);
    A->cluster_sz.upload(csz);
    A->rowptr.upload(std::vector<int64_t>(rowptr, rowptr + nclusters + 1));
    A->rowptr_val.upload(std::vector<int64_t>(rowptr_val, rowptr_val + nclusters + 1));
    A->colids.upload(std::vector<int32_t>(colids, colids + nnzc));
    A->values.upload_raw(values, (size_t)nnzv * value_bytes(value_type));
    A->perm.upload(perm);
    A->offsets.upload(offsets);
    A->cid.upload(cid);
    A->lrow.upload(lrow);
    *out = A.release();
    return SPOPS_SUCCESS;
  });
}
