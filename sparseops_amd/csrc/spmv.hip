// SpMV on MI355X (gfx950): CSR, SELL-C-sigma / SELL-C-R, ELL and BSR.
//
// Semantics follow the reference kernels (y = alpha*A*x + beta*y):
//   CSR   src/spmv_csr.cpp:34-185        ELL  src/spmv_ell.cpp:17-247
//   SELL  src/spmv_sell_c_sigma.cpp:18-211 (SELL-C-R: src/spmv_sell_c_R.cpp, sigma = n)
//   BSR   src/spmv_bsr.cpp:18-233
// with two deliberate differences: beta == 0 never reads y, and padding slots are never
// dereferenced (the reference reads x[(size_t)-1] for SELL padding and x past num_cols
// for BSR).
//
// Device layouts (HBM), all built once at matrix creation:
//   CSR  : rowptr (int32, or int64 when nnz >= 2^31), col int32, values V; a launch plan of
//          wave tasks (<= 512-entry blocks of consecutive rows with <= 64 rows, or one
//          medium row) and a list of long rows (> 2048 entries), each one workgroup.
//   SELL : rows grouped in chunks of C sorted rows.  Inside a chunk the entries of row l
//          are stored in groups of 4 consecutive entries, group q of row l at
//          base + q*C + l  (so one half-wave loads C*16 contiguous bytes with 128-bit
//          loads per lane).  Each sorted row carries (original row, length); loads past a
//          row's length are masked, so padding is never fetched.  Rows much longer than
//          their chunk's median are moved to an overflow CSR and handled by one wave or
//          one workgroup each, which keeps every chunk's wave short.
//   ELL  : the SELL layout with sigma = 1 (no reordering), C = 32.
//   BSR  : the reference layout (row_ptr, block_colindex, row-major bR x bC blocks).
// The streams of A (indices, values) use non-temporal loads so that L2 keeps x.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "common.hpp"
#include "spmv_internal.hpp"

namespace spops {

// ---------------------------------------------------------------------------------------------
// value-type traits
// ---------------------------------------------------------------------------------------------
template <int VT> struct VT_;
template <> struct VT_<SPOPS_VALUE_F64> {
  using V = double; using X = double; using Acc = double;
  static constexpr bool kComp = false;
  typedef double G __attribute__((ext_vector_type(4)));    // a 4-entry value group
  static __device__ __forceinline__ double cvt(double v) { return v; }
};
template <> struct VT_<SPOPS_VALUE_F32> {
  using V = float; using X = float; using Acc = float;
  static constexpr bool kComp = true;
  typedef float G __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ float cvt(float v) { return v; }
};
template <> struct VT_<SPOPS_VALUE_BF16> {
  using V = uint16_t; using X = float; using Acc = float;
  static constexpr bool kComp = true;
  typedef uint16_t G __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ float cvt(uint16_t v) { return bf16_to_float(v); }
};

typedef int v4i __attribute__((ext_vector_type(4)));

template <typename X, typename Acc>
__device__ __forceinline__ void store_y(X *y, long row, Acc alpha, Acc beta, Acc sum) {
  Acc r = alpha * sum;
  if (beta != Acc(0)) r = r + beta * (Acc)y[row];
  y[row] = (X)r;
}

// Sum over entries [0, len) of one row whose col/val start at the given pointers: lanes of
// one wave stride by 64, four entries per lane in flight, every load a buffer load (masked
// lanes fetch nothing), gathers of x through the x resource.
template <int VT>
__device__ __forceinline__ void row_sum_wave(const int *col, const typename VT_<VT>::V *val, int64_t len, rsrc_t rx,
                                             Accum<typename VT_<VT>::Acc, VT_<VT>::kComp> &a) {
  using T = VT_<VT>;
  constexpr uint32_t sV = sizeof(typename T::V), sX = sizeof(typename T::X);
  const int lane = lane_id();
  for (int64_t seg = 0; seg < len; seg += kRowSegment) {
    const int n = (int)((len - seg) < kRowSegment ? (len - seg) : kRowSegment);
    const rsrc_t rc = make_rsrc(col + seg, (uint64_t)n * 4), rv = make_rsrc(val + seg, (uint64_t)n * sV);
    for (int p = lane; p < n; p += 4 * kWave) {
      int c[4];
      typename T::V v[4];
      typename T::X xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = p + u * kWave;
        c[u] = bload<int, kAuxNT>(rc, e < n ? (uint32_t)e * 4u : kOOB);
        v[u] = bload<typename T::V, kAuxNT>(rv, e < n ? (uint32_t)e * sV : kOOB);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = bload<typename T::X>(rx, (p + u * kWave < n) ? (uint32_t)c[u] * sX : kOOB);
#pragma unroll
      for (int u = 0; u < 4; ++u) a.fma(T::cvt(v[u]), (typename T::Acc)xv[u]);
    }
  }
}

// Whole-workgroup sum of one long row; result valid in thread 0.
template <int VT>
__device__ __forceinline__ typename VT_<VT>::Acc row_sum_block(const int *col, const typename VT_<VT>::V *val,
                                                              int64_t len, rsrc_t rx,
                                                              typename VT_<VT>::Acc *scratch /* 8 entries of LDS */) {
  using T = VT_<VT>;
  constexpr uint32_t sV = sizeof(typename T::V), sX = sizeof(typename T::X);
  Accum<typename T::Acc, T::kComp> a;
  for (int64_t seg = 0; seg < len; seg += kRowSegment) {
    const int n = (int)((len - seg) < kRowSegment ? (len - seg) : kRowSegment);
    const rsrc_t rc = make_rsrc(col + seg, (uint64_t)n * 4), rv = make_rsrc(val + seg, (uint64_t)n * sV);
    for (int p = (int)threadIdx.x; p < n; p += 4 * kBlock) {
      int c[4];
      typename T::V v[4];
      typename T::X xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = p + u * kBlock;
        c[u] = bload<int, kAuxNT>(rc, e < n ? (uint32_t)e * 4u : kOOB);
        v[u] = bload<typename T::V, kAuxNT>(rv, e < n ? (uint32_t)e * sV : kOOB);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = bload<typename T::X>(rx, (p + u * kBlock < n) ? (uint32_t)c[u] * sX : kOOB);
#pragma unroll
      for (int u = 0; u < 4; ++u) a.fma(T::cvt(v[u]), (typename T::Acc)xv[u]);
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) a.merge_xor(m);
  if (lane_id() == 0) {
    scratch[wave_id()] = a.s;
    if constexpr (T::kComp) scratch[4 + wave_id()] = a.c;
  }
  __syncthreads();
  typename T::Acc r = 0;
  if (threadIdx.x == 0) {
    Accum<typename T::Acc, T::kComp> t;
    for (int w = 0; w < kBlock / kWave; ++w) {
      t.add(scratch[w]);
      if constexpr (T::kComp) t.add(scratch[4 + w]);
    }
    r = t.value();
  }
  return r;
}

// ---------------------------------------------------------------------------------------------
// CSR
// ---------------------------------------------------------------------------------------------
// grid = nlong workgroups (one long row each) + ceil(ntasks / 4) workgroups of 4 wave tasks.
template <int VT, typename RP>
__global__ __launch_bounds__(kBlock) void k_csr_spmv(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                     const typename VT_<VT>::V *__restrict__ val,
                                                     const typename VT_<VT>::X *__restrict__ x, int n_cols,
                                                     typename VT_<VT>::X *__restrict__ y,
                                                     typename VT_<VT>::Acc alpha, typename VT_<VT>::Acc beta,
                                                     const int2 *__restrict__ tasks, int ntasks,
                                                     const int *__restrict__ long_rows, int nlong) {
  using T = VT_<VT>;
  using Acc = typename T::Acc;
  constexpr uint32_t sV = sizeof(typename T::V), sX = sizeof(typename T::X);
  __shared__ Acc lds[kBlock / kWave][kCsrStreamNnz];
  const rsrc_t rx = make_rsrc(x, (uint64_t)n_cols * sX);

  if ((int)blockIdx.x < nlong) {
    const int row = long_rows[blockIdx.x];
    const RP p0 = rowptr[row], p1 = rowptr[row + 1];
    Acc r = row_sum_block<VT>(col + p0, val + p0, (int64_t)(p1 - p0), rx, &lds[0][0]);
    if (threadIdx.x == 0) store_y(y, row, alpha, beta, r);
    return;
  }
  const int t = __builtin_amdgcn_readfirstlane(((int)blockIdx.x - nlong) * (kBlock / kWave) + wave_id());
  if (t >= ntasks) return;
  const int2 tk = tasks[t];
  const int r0 = tk.x, r1 = tk.y;
  const RP p0 = rowptr[r0], p1 = rowptr[r1];
  const int lane = lane_id();

  if (r1 - r0 == 1 && p1 - p0 > kCsrStreamNnz) {        // one medium row: wave-per-row
    Accum<Acc, T::kComp> a;
    row_sum_wave<VT>(col + p0, val + p0, (int64_t)(p1 - p0), rx, a);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) a.merge_xor(m);
    if (lane == 0) store_y(y, r0, alpha, beta, a.value());
    return;
  }

  // stream block: stage the products of <= 512 consecutive entries in LDS, then reduce
  // each row with 64 / 2^ceil(log2 R) lanes (R = rows in the block).
  Acc *sh = lds[wave_id()];
  const int cnt = (int)(p1 - p0);
  const rsrc_t rc = make_rsrc(col + p0, (uint64_t)cnt * 4), rv = make_rsrc(val + p0, (uint64_t)cnt * sV);
  constexpr int K = kCsrStreamNnz / kWave;
  int c[K];
  typename T::V v[K];
  typename T::X xv[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int e = lane + i * kWave;
    c[i] = bload<int, kAuxNT>(rc, e < cnt ? (uint32_t)e * 4u : kOOB);
    v[i] = bload<typename T::V, kAuxNT>(rv, e < cnt ? (uint32_t)e * sV : kOOB);
  }
#pragma unroll
  for (int i = 0; i < K; ++i) xv[i] = bload<typename T::X>(rx, (lane + i * kWave < cnt) ? (uint32_t)c[i] * sX : kOOB);
#pragma unroll
  for (int i = 0; i < K; ++i) sh[lane + i * kWave] = T::cvt(v[i]) * (Acc)xv[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const int R = r1 - r0;
  const int lg = R <= 1 ? 6 : 6 - (32 - __clz(R - 1));
  const int lpr = 1 << lg;
  const int g = lane >> lg, sub = lane & (lpr - 1);
  Accum<Acc, T::kComp> a;
  if (g < R) {
    const int b = (int)(rowptr[r0 + g] - p0), e = (int)(rowptr[r0 + g + 1] - p0);
    for (int j = b + sub; j < e; j += lpr) a.add(sh[j]);
  }
  for (int m = lpr >> 1; m > 0; m >>= 1) a.merge_xor(m);
  if (g < R && sub == 0) store_y(y, r0 + g, alpha, beta, a.value());
}

// ---------------------------------------------------------------------------------------------
// SELL-C-sigma (also ELL).  grid = n_ovf_long workgroups + ceil((n_ovf_vec + nchunks)/4).
// ---------------------------------------------------------------------------------------------
template <int VT, int C>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(const v4i *__restrict__ gcol,
                                                      const typename VT_<VT>::G *__restrict__ gval,
                                                      const int2 *__restrict__ rinfo, const long *__restrict__ cbase,
                                                      const int *__restrict__ cwidth, int nchunks, int n_rows,
                                                      const long *__restrict__ optr, const int *__restrict__ ocol,
                                                      const typename VT_<VT>::V *__restrict__ oval,
                                                      const int *__restrict__ orow, int n_olong, int n_ovec,
                                                      const typename VT_<VT>::X *__restrict__ x, int n_cols,
                                                      typename VT_<VT>::X *__restrict__ y,
                                                      typename VT_<VT>::Acc alpha, typename VT_<VT>::Acc beta) {
  using T = VT_<VT>;
  using Acc = typename T::Acc;
  constexpr uint32_t sX = sizeof(typename T::X), sG = sizeof(typename T::G);
  __shared__ Acc lds[8];
  const rsrc_t rx = make_rsrc(x, (uint64_t)n_cols * sX);
  if ((int)blockIdx.x < n_olong) {
    const int o = blockIdx.x;
    const long p0 = optr[o];
    Acc r = row_sum_block<VT>(ocol + p0, oval + p0, optr[o + 1] - p0, rx, lds);
    if (threadIdx.x == 0) store_y(y, orow[o], alpha, beta, r);
    return;
  }
  const int t = __builtin_amdgcn_readfirstlane(((int)blockIdx.x - n_olong) * (kBlock / kWave) + wave_id());
  const int lane = lane_id();
  if (t < n_ovec) {
    const int o = n_olong + t;
    const long p0 = optr[o];
    Accum<Acc, T::kComp> a;
    row_sum_wave<VT>(ocol + p0, oval + p0, optr[o + 1] - p0, rx, a);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) a.merge_xor(m);
    if (lane == 0) store_y(y, orow[o], alpha, beta, a.value());
    return;
  }
  const int ch = t - n_ovec;
  if (ch >= nchunks) return;
  constexpr int H = kWave / C;              // lanes per row
  const int l = lane % C, h = lane / C;
  const long base = cbase[ch];
  const int W = cwidth[ch];
  const rsrc_t rc = make_rsrc(gcol + base, (uint64_t)W * C * 16), rv = make_rsrc(gval + base, (uint64_t)W * C * sG);
  const int gr = ch * C + l;
  const int2 ri = gr < n_rows ? rinfo[gr] : make_int2(-1, -1);
  const int len = ri.y;                     // -1: padding row or row moved to the overflow
  Accum<Acc, T::kComp> a;
  constexpr int U = (VT == SPOPS_VALUE_F64) ? 2 : 4;
  for (int q0 = h; q0 < W; q0 += H * U) {
    v4i cc[U];
    typename T::G vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * H;
      const bool act = 4 * q < len;
      const uint32_t slot = (uint32_t)(q * C + l);
      cc[u] = bload<v4i, kAuxNT>(rc, act ? slot * 16u : kOOB);
      if constexpr (sG == 32) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 lo = bload<d2, kAuxNT>(rv, act ? slot * 32u : kOOB);
        const d2 hi = bload<d2, kAuxNT>(rv, act ? slot * 32u + 16u : kOOB);
        vv[u] = typename T::G{lo.x, lo.y, hi.x, hi.y};
      } else {
        vv[u] = bload<typename T::G, kAuxNT>(rv, act ? slot * sG : kOOB);
      }
    }
    typename T::X xv[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        xv[u][e] = bload<typename T::X>(rx, (4 * (q0 + u * H) + e < len) ? (uint32_t)cc[u][e] * sX : kOOB);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) a.fma(T::cvt(vv[u][e]), (Acc)xv[u][e]);
  }
#pragma unroll
  for (int m = C; m < kWave; m <<= 1) a.merge_xor(m);
  if (h == 0 && len >= 0) store_y(y, ri.x, alpha, beta, a.value());
}

// ---------------------------------------------------------------------------------------------
// BSR: one wave per block row; products of the bR x bC block staged in LDS per lane slot.
// ---------------------------------------------------------------------------------------------
template <int VT>
__global__ __launch_bounds__(kBlock) void k_bsr_spmv(const int *__restrict__ brp, const int *__restrict__ bcol,
                                                     const typename VT_<VT>::V *__restrict__ bdata, int mb, int bR,
                                                     int bC, int n_rows, int n_cols,
                                                     const typename VT_<VT>::X *__restrict__ x,
                                                     typename VT_<VT>::X *__restrict__ y,
                                                     typename VT_<VT>::Acc alpha, typename VT_<VT>::Acc beta) {
  using T = VT_<VT>;
  using Acc = typename T::Acc;
  constexpr uint32_t sV = sizeof(typename T::V), sX = sizeof(typename T::X);
  __shared__ Acc lds[kBlock / kWave][kBsrMaxBlock];
  const int i = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (kBlock / kWave) + wave_id());
  if (i >= mb) return;
  const int lane = lane_id();
  const int bn = bR * bC;
  const rsrc_t rx = make_rsrc(x, (uint64_t)n_cols * sX);
  Accum<Acc, T::kComp> a[kBsrMaxBlock / kWave];
  const int b0 = brp[i], b1 = brp[i + 1];
  const rsrc_t rd = make_rsrc(bdata + (long)b0 * bn, (uint64_t)(b1 - b0) * bn * sV);
  for (int b = b0; b < b1; ++b) {
    const long xb = (long)bcol[b] * bC;
    typename T::V v[kBsrMaxBlock / kWave];
    typename T::X xv[kBsrMaxBlock / kWave];
#pragma unroll
    for (int k = 0; k < kBsrMaxBlock / kWave; ++k) {
      const int e = lane + k * kWave;
      v[k] = bload<typename T::V, kAuxNT>(rd, e < bn ? (uint32_t)((b - b0) * bn + e) * sV : kOOB);
      const long xi = xb + e % bC;
      xv[k] = bload<typename T::X>(rx, (e < bn && xi < n_cols) ? (uint32_t)xi * sX : kOOB);
    }
#pragma unroll
    for (int k = 0; k < kBsrMaxBlock / kWave; ++k) a[k].fma(T::cvt(v[k]), (Acc)xv[k]);
  }
  Acc *sh = lds[wave_id()];
#pragma unroll
  for (int k = 0; k < kBsrMaxBlock / kWave; ++k) sh[lane + k * kWave] = a[k].value();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < bR) {
    const long row = (long)i * bR + lane;
    if (row < n_rows) {
      Accum<Acc, T::kComp> s;
      for (int c = 0; c < bC; ++c) s.add(sh[lane * bC + c]);
      store_y(y, row, alpha, beta, s.value());
    }
  }
}

// ---------------------------------------------------------------------------------------------
// host: plans and layouts
// ---------------------------------------------------------------------------------------------
namespace {

template <typename I>
std::vector<int64_t> widen(const void *p, int64_t n) {
  const I *q = static_cast<const I *>(p);
  return std::vector<int64_t>(q, q + n);
}

std::vector<int64_t> host_index(const void *p, int64_t n, int index_type) {
  SPOPS_REQUIRE(index_type == SPOPS_INDEX_I32 || index_type == SPOPS_INDEX_I64, SPOPS_ERROR_INVALID_VALUE,
                "index_type must be SPOPS_INDEX_I32 or SPOPS_INDEX_I64");
  SPOPS_REQUIRE(n == 0 || p != nullptr, SPOPS_ERROR_INVALID_VALUE, "null index array");
  return index_type == SPOPS_INDEX_I32 ? widen<int32_t>(p, n) : widen<int64_t>(p, n);
}

double host_value(const void *p, int in_vt, int64_t i) {
  return in_vt == SPOPS_VALUE_F64 ? static_cast<const double *>(p)[i] : (double)static_cast<const float *>(p)[i];
}

// Convert n host values (double or float) to the storage type, bytes appended to `out`.
void pack_values(std::vector<uint8_t> &out, int vt, double v) {
  if (vt == SPOPS_VALUE_F64) {
    uint8_t b[8]; std::memcpy(b, &v, 8); out.insert(out.end(), b, b + 8);
  } else if (vt == SPOPS_VALUE_F32) {
    float f = (float)v; uint8_t b[4]; std::memcpy(b, &f, 4); out.insert(out.end(), b, b + 4);
  } else {
    uint16_t h = double_to_bf16_rne(v); uint8_t b[2]; std::memcpy(b, &h, 2); out.insert(out.end(), b, b + 2);
  }
}

void check_csr(int64_t n_rows, int64_t n_cols, int64_t nnz, const std::vector<int64_t> &rp,
               const std::vector<int64_t> &ci) {
  SPOPS_REQUIRE(n_rows >= 0 && n_cols >= 0 && nnz >= 0, SPOPS_ERROR_INVALID_VALUE, "negative matrix size");
  SPOPS_REQUIRE((uint64_t)n_cols * 8 <= kMaxBufBytes, SPOPS_ERROR_LIMIT,
                "x must fit one 2 GiB buffer resource (n_cols <= 268435454)");
  SPOPS_REQUIRE(n_rows < (int64_t)INT32_MAX && n_cols < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT,
                "rows and columns must be < 2^31 (got %lld x %lld)", (long long)n_rows, (long long)n_cols);
  SPOPS_REQUIRE(rp[0] == 0 && rp[n_rows] == nnz, SPOPS_ERROR_INVALID_VALUE,
                "row_offset must start at 0 and end at nnz");
  for (int64_t r = 0; r < n_rows; ++r)
    SPOPS_REQUIRE(rp[r + 1] >= rp[r], SPOPS_ERROR_INVALID_VALUE, "row_offset not monotone at row %lld", (long long)r);
  for (int64_t j = 0; j < nnz; ++j)
    SPOPS_REQUIRE(ci[j] >= 0 && ci[j] < n_cols, SPOPS_ERROR_INVALID_VALUE,
                  "column index %lld out of range at entry %lld", (long long)ci[j], (long long)j);
}

}  // namespace

// CSR plan: consecutive-row stream blocks of <= kCsrStreamNnz entries and <= 64 rows; a
// row longer than kCsrStreamNnz is a wave task of its own, longer than kCsrLongRow a
// workgroup task.
void build_csr_plan(const std::vector<int64_t> &rp, int64_t n_rows, std::vector<int2> &tasks, std::vector<int> &longs) {
  tasks.clear();
  longs.clear();
  int64_t r = 0;
  while (r < n_rows) {
    int64_t len = rp[r + 1] - rp[r];
    if (len > kCsrLongRow) { longs.push_back((int)r); ++r; continue; }
    if (len > kCsrStreamNnz) { tasks.push_back(make_int2((int)r, (int)r + 1)); ++r; continue; }
    int64_t r1 = r, acc = 0;
    while (r1 < n_rows && r1 - r < kWave) {
      int64_t l = rp[r1 + 1] - rp[r1];
      if (l > kCsrStreamNnz || acc + l > kCsrStreamNnz) break;
      acc += l;
      ++r1;
    }
    tasks.push_back(make_int2((int)r, (int)r1));
    r = r1;
  }
}

// Decide which rows of a device chunk leave for the overflow CSR: rows longer than
// max(kSellOvfMin, 4 x (median length of the chunk's rows)).
static void sell_split_chunk(const int64_t *len, int m, std::vector<char> &out) {
  std::vector<int64_t> s(len, len + m);
  std::nth_element(s.begin(), s.begin() + m / 2, s.end());
  int64_t med = s[m / 2];
  int64_t thr = std::max<int64_t>(kSellOvfMin, 4 * std::max<int64_t>(med, 4));
  out.assign(m, 0);
  for (int l = 0; l < m; ++l) out[l] = len[l] > thr;
}

// Row access for the SELL builder: entries of the row at sorted position g.
struct RowSource {
  virtual ~RowSource() = default;
  virtual int64_t length(int64_t g) const = 0;
  virtual int64_t orig_row(int64_t g) const = 0;
  virtual void entry(int64_t g, int64_t k, int64_t &col, double &val) const = 0;
};

void build_sell_layout(spops_matrix_s &M, const RowSource &src, int vt, int Cd) {
  const int64_t n = M.n_rows;
  const int64_t nch = (n + Cd - 1) / Cd;
  std::vector<int2> rinfo(n);
  std::vector<long> cbase(nch);
  std::vector<int> cwidth(nch);
  std::vector<int64_t> ovf_rows;                       // sorted positions moved to overflow
  std::vector<char> ovf_flag;
  int64_t groups = 0;
  std::vector<int64_t> lens(Cd);
  for (int64_t c = 0; c < nch; ++c) {
    int m = (int)std::min<int64_t>(Cd, n - c * Cd);
    for (int l = 0; l < m; ++l) lens[l] = src.length(c * Cd + l);
    sell_split_chunk(lens.data(), m, ovf_flag);
    int64_t w = 0;
    for (int l = 0; l < m; ++l) {
      int64_t g = c * Cd + l;
      if (ovf_flag[l]) { ovf_rows.push_back(g); rinfo[g] = make_int2((int)src.orig_row(g), -1); }
      else { rinfo[g] = make_int2((int)src.orig_row(g), (int)lens[l]); w = std::max(w, (lens[l] + 3) / 4); }
    }
    cbase[c] = groups;
    cwidth[c] = (int)w;
    groups += w * Cd;
  }
  // group arrays
  const size_t vb = value_bytes(vt);
  std::vector<int32_t> gcol((size_t)groups * 4, 0);
  std::vector<uint8_t> gval((size_t)groups * 4 * vb, 0);
  auto put_val = [&](size_t slot, double v) {
    if (vt == SPOPS_VALUE_F64) std::memcpy(&gval[slot * 8], &v, 8);
    else if (vt == SPOPS_VALUE_F32) { float f = (float)v; std::memcpy(&gval[slot * 4], &f, 4); }
    else { uint16_t h = double_to_bf16_rne(v); std::memcpy(&gval[slot * 2], &h, 2); }
  };

  for (int64_t c = 0; c < nch; ++c) {
    int m = (int)std::min<int64_t>(Cd, n - c * Cd);
    for (int l = 0; l < m; ++l) {
      int64_t g = c * Cd + l;
      int len = rinfo[g].y;
      for (int64_t k = 0; k < len; ++k) {
        int64_t col; double v;
        src.entry(g, k, col, v);
        size_t slot = (size_t)(cbase[c] + (k / 4) * Cd + l) * 4 + (k % 4);
        gcol[slot] = (int32_t)col;
        put_val(slot, v);
      }
    }
  }
  // overflow CSR: workgroup rows (> kCsrLongRow) first, then wave rows
  std::vector<int64_t> olong, ovec;
  for (int64_t g : ovf_rows) (src.length(g) > kCsrLongRow ? olong : ovec).push_back(g);
  std::vector<long> optr(1, 0);
  std::vector<int32_t> ocol, orow;
  std::vector<uint8_t> oval;
  for (auto *lst : {&olong, &ovec})
    for (int64_t g : *lst) {
      int64_t len = src.length(g);
      for (int64_t k = 0; k < len; ++k) {
        int64_t col; double v;
        src.entry(g, k, col, v);
        ocol.push_back((int32_t)col);
        pack_values(oval, vt, v);
      }
      optr.push_back(optr.back() + len);
      orow.push_back((int32_t)src.orig_row(g));
    }
  M.sell.gcol.upload(gcol);
  M.sell.gval.upload(gval);
  M.sell.rinfo.upload(rinfo);
  M.sell.cbase.upload(cbase);
  M.sell.cwidth.upload(cwidth);
  M.sell.optr.upload(optr);
  M.sell.ocol.upload(ocol);
  M.sell.oval.upload(oval);
  M.sell.orow.upload(orow);
  M.sell.nchunks = nch;
  M.sell.Cd = Cd;
  M.sell.n_olong = (int)olong.size();
  M.sell.n_ovec = (int)ovec.size();
  M.slots = groups * 4 + (int64_t)ocol.size();
  M.dev_bytes = M.sell.gcol.bytes + M.sell.gval.bytes + M.sell.rinfo.bytes + M.sell.cbase.bytes + M.sell.cwidth.bytes +
                M.sell.optr.bytes + M.sell.ocol.bytes + M.sell.oval.bytes + M.sell.orow.bytes;
}

// ---------------------------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------------------------
template <int VT>
static void launch_spmv_vt(const spops_matrix_s &M, double alpha, const void *x, double beta, void *y, hipStream_t s) {
  using T = VT_<VT>;
  using Acc = typename T::Acc;
  const auto *xx = static_cast<const typename T::X *>(x);
  auto *yy = static_cast<typename T::X *>(y);
  if (M.format == SPOPS_FORMAT_CSR) {
    const auto &P = M.csr;
    int grid = P.nlong + (P.ntasks + 3) / 4;
    if (grid == 0) return;
    if (P.rowptr64)
      hipLaunchKernelGGL((k_csr_spmv<VT, long>), dim3(grid), dim3(kBlock), 0, s, P.rowptr.as<long>(), P.col.as<int>(),
                         P.val.as<typename T::V>(), xx, (int)M.n_cols, yy, (Acc)alpha, (Acc)beta, P.tasks.as<int2>(), P.ntasks,
                         P.longs.as<int>(), P.nlong);
    else
      hipLaunchKernelGGL((k_csr_spmv<VT, int>), dim3(grid), dim3(kBlock), 0, s, P.rowptr.as<int>(), P.col.as<int>(),
                         P.val.as<typename T::V>(), xx, (int)M.n_cols, yy, (Acc)alpha, (Acc)beta, P.tasks.as<int2>(), P.ntasks,
                         P.longs.as<int>(), P.nlong);
  } else if (M.format == SPOPS_FORMAT_SELL || M.format == SPOPS_FORMAT_ELL) {
    const auto &S = M.sell;
    int grid = S.n_olong + (int)((S.n_ovec + S.nchunks + 3) / 4);
    if (grid == 0) return;
#define SPOPS_SELL_LAUNCH(CC)                                                                                   \
  hipLaunchKernelGGL((k_sell_spmv<VT, CC>), dim3(grid), dim3(kBlock), 0, s, S.gcol.as<v4i>(),                  \
                     S.gval.as<typename T::G>(), S.rinfo.as<int2>(), S.cbase.as<long>(), S.cwidth.as<int>(),   \
                     (int)S.nchunks, (int)M.n_rows, S.optr.as<long>(), S.ocol.as<int>(),                       \
                     S.oval.as<typename T::V>(), S.orow.as<int>(), S.n_olong, S.n_ovec, xx, (int)M.n_cols, yy,     \
                     (Acc)alpha,                                                                             \
                     (Acc)beta)
    switch (S.Cd) {
      case 4: SPOPS_SELL_LAUNCH(4); break;
      case 8: SPOPS_SELL_LAUNCH(8); break;
      case 16: SPOPS_SELL_LAUNCH(16); break;
      case 32: SPOPS_SELL_LAUNCH(32); break;
      default: SPOPS_SELL_LAUNCH(64); break;
    }
#undef SPOPS_SELL_LAUNCH
  } else {
    const auto &B = M.bsr;
    int grid = (int)((B.mb + 3) / 4);
    if (grid == 0) return;
    hipLaunchKernelGGL((k_bsr_spmv<VT>), dim3(grid), dim3(kBlock), 0, s, B.brp.as<int>(), B.bcol.as<int>(),
                       B.data.as<typename T::V>(), (int)B.mb, (int)B.bR, (int)B.bC, (int)M.n_rows, (int)M.n_cols,
                       xx, yy, (Acc)alpha, (Acc)beta);
  }
  SPOPS_HIP_CHECK(hipGetLastError());
}

void launch_spmv(const spops_matrix_s &M, double alpha, const void *x, double beta, void *y, hipStream_t s) {
  switch (M.value_type) {
    case SPOPS_VALUE_F64: launch_spmv_vt<SPOPS_VALUE_F64>(M, alpha, x, beta, y, s); break;
    case SPOPS_VALUE_F32: launch_spmv_vt<SPOPS_VALUE_F32>(M, alpha, x, beta, y, s); break;
    default: launch_spmv_vt<SPOPS_VALUE_BF16>(M, alpha, x, beta, y, s); break;
  }
}

// algorithmic bytes of one SpMV: (n+1)*sI + nnz*(sI+sV) + n_cols*sX + n_rows*sX
// (+ n_rows*sX when beta != 0 is counted by the caller), plus n*4 + chunks*4 for SELL.
int64_t spmv_algo_bytes(const spops_matrix_s &M, int64_t sI) {
  int64_t sV = (int64_t)value_bytes(M.value_type), sX = (int64_t)xy_bytes(M.value_type);
  int64_t b = (M.n_rows + 1) * sI + M.nnz * (sI + sV) + M.n_cols * sX + M.n_rows * sX;
  if (M.format == SPOPS_FORMAT_SELL) b += M.n_rows * 4 + M.ref_nchunks * 4;
  return b;
}

}  // namespace spops

// =============================================================================================
// C ABI
// =============================================================================================
using namespace spops;

namespace {

int check_vt(int in_vt, int vt) {
  SPOPS_REQUIRE(in_vt == SPOPS_VALUE_F64 || in_vt == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                "input values must be F64 or F32");
  SPOPS_REQUIRE(vt == SPOPS_VALUE_F64 || vt == SPOPS_VALUE_F32 || vt == SPOPS_VALUE_BF16, SPOPS_ERROR_INVALID_VALUE,
                "value_type must be F64, F32 or BF16");
  return 0;
}

struct CsrRows : RowSource {
  const std::vector<int64_t> &rp, &ci, &perm;
  const void *vals;
  int in_vt;
  CsrRows(const std::vector<int64_t> &r, const std::vector<int64_t> &c, const std::vector<int64_t> &p, const void *v, int t)
      : rp(r), ci(c), perm(p), vals(v), in_vt(t) {}
  int64_t length(int64_t g) const override { return rp[perm[g] + 1] - rp[perm[g]]; }
  int64_t orig_row(int64_t g) const override { return perm[g]; }
  void entry(int64_t g, int64_t k, int64_t &col, double &val) const override {
    int64_t j = rp[perm[g]] + k;
    col = ci[j];
    val = host_value(vals, in_vt, j);
  }
};

int device_chunk_height(int64_t C) {
  if (C == 4 || C == 8 || C == 16 || C == 32 || C == 64) return (int)C;
  return 32;
}

// the reference's per-slice order: sort by (nnz, row) descending inside each sigma-slice
std::vector<int64_t> sell_reorder(const std::vector<int64_t> &rp, int64_t n, int64_t sigma) {
  std::vector<int64_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int64_t lo = 0; lo < n; lo += sigma) {
    int64_t hi = std::min(n, lo + sigma);
    std::sort(perm.begin() + lo, perm.begin() + hi, [&](int64_t a, int64_t b) {
      int64_t la = rp[a + 1] - rp[a], lb = rp[b + 1] - rp[b];
      if (la != lb) return la > lb;
      return a > b;
    });
  }
  return perm;
}

// the reference's chunk_len: max nnz of each C sorted rows, rounded to 256/8/sizeof(V)
std::vector<int32_t> sell_ref_chunk_len(const std::vector<int64_t> &rp, const std::vector<int64_t> &perm, int64_t n,
                                        int64_t C, int64_t align) {
  int64_t nch = (n + C - 1) / C;
  std::vector<int32_t> cl(nch, 0);
  for (int64_t g = 0; g < n; ++g) cl[g / C] = std::max<int32_t>(cl[g / C], (int32_t)(rp[perm[g] + 1] - rp[perm[g]]));
  for (auto &v : cl) v = (int32_t)((v + align - 1) / align * align);
  return cl;
}

}  // namespace

extern "C" int spops_csr_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                                const void *rowptr, const void *col, const void *values, int in_vt, int vt) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    std::vector<int64_t> rp = host_index(rowptr, n_rows + 1, index_type);
    std::vector<int64_t> ci = host_index(col, nnz, index_type);
    check_csr(n_rows, n_cols, nnz, rp, ci);
    auto M = std::make_unique<spops_matrix_s>();
    M->format = SPOPS_FORMAT_CSR;
    M->value_type = vt;
    M->index_type = index_type;
    M->n_rows = n_rows; M->n_cols = n_cols; M->nnz = nnz;
    auto &P = M->csr;
    P.rowptr64 = nnz >= (int64_t)INT32_MAX;
    if (P.rowptr64) P.rowptr.upload(rp);
    else P.rowptr.upload(std::vector<int32_t>(rp.begin(), rp.end()));
    P.col.upload(std::vector<int32_t>(ci.begin(), ci.end()));
    std::vector<uint8_t> v;
    v.reserve((size_t)nnz * value_bytes(vt));
    for (int64_t j = 0; j < nnz; ++j) pack_values(v, vt, host_value(values, in_vt, j));
    P.val.upload(v);
    std::vector<int2> tasks;
    std::vector<int> longs;
    build_csr_plan(rp, n_rows, tasks, longs);
    P.tasks.upload(tasks);
    P.longs.upload(longs);
    P.ntasks = (int)tasks.size();
    P.nlong = (int)longs.size();
    M->slots = nnz;
    M->dev_bytes = P.rowptr.bytes + P.col.bytes + P.val.bytes + P.tasks.bytes + P.longs.bytes;
    M->algo_bytes = spmv_algo_bytes(*M, index_type == SPOPS_INDEX_I32 ? 4 : 8);
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}

static int sell_create_common(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                              const std::vector<int64_t> &rp, const std::vector<int64_t> &ci, const void *values,
                              int in_vt, int vt, int64_t sigma, int64_t C, int format) {
  auto M = std::make_unique<spops_matrix_s>();
  M->format = format;
  M->value_type = vt;
  M->index_type = index_type;
  M->n_rows = n_rows; M->n_cols = n_cols; M->nnz = nnz;
  std::vector<int64_t> perm = sell_reorder(rp, n_rows, sigma);
  int Cd = format == SPOPS_FORMAT_ELL ? 32 : device_chunk_height(C);
  CsrRows src(rp, ci, perm, values, in_vt);
  build_sell_layout(*M, src, vt, Cd);
  M->sell.C = C;
  M->sell.sigma = sigma;
  M->ref_nchunks = (n_rows + C - 1) / C;
  M->sell.ref_reorder.assign(perm.begin(), perm.end());
  int64_t align = 256 / 8 / (int64_t)(vt == SPOPS_VALUE_F64 ? 8 : 4);
  M->sell.ref_chunk_len = sell_ref_chunk_len(rp, perm, n_rows, C, align);
  M->algo_bytes = spmv_algo_bytes(*M, index_type == SPOPS_INDEX_I32 ? 4 : 8);
  *A = M.release();
  return SPOPS_SUCCESS;
}

extern "C" int spops_sell_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                                 const void *rowptr, const void *col, const void *values, int in_vt, int vt,
                                 int64_t sigma, int64_t C) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    SPOPS_REQUIRE(C > 0 && sigma > 0, SPOPS_ERROR_INVALID_VALUE, "sigma and C must be positive");
    SPOPS_REQUIRE(sigma % C == 0, SPOPS_ERROR_INVALID_VALUE, "Sigma must be multiple divided by c (sigma=%lld C=%lld)",
                  (long long)sigma, (long long)C);
    std::vector<int64_t> rp = host_index(rowptr, n_rows + 1, index_type);
    std::vector<int64_t> ci = host_index(col, nnz, index_type);
    check_csr(n_rows, n_cols, nnz, rp, ci);
    return sell_create_common(A, n_rows, n_cols, nnz, index_type, rp, ci, values, in_vt, vt, sigma, C,
                              SPOPS_FORMAT_SELL);
  });
}

extern "C" int spops_ell_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                                const void *rowptr, const void *col, const void *values, int in_vt, int vt) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    std::vector<int64_t> rp = host_index(rowptr, n_rows + 1, index_type);
    std::vector<int64_t> ci = host_index(col, nnz, index_type);
    check_csr(n_rows, n_cols, nnz, rp, ci);
    int r = sell_create_common(A, n_rows, n_cols, nnz, index_type, rp, ci, values, in_vt, vt, 1, 1, SPOPS_FORMAT_ELL);
    int64_t w = 0;
    for (int64_t i = 0; i < n_rows; ++i) w = std::max(w, rp[i + 1] - rp[i]);
    (*A)->ell_width = w;
    return r;
  });
}

namespace {
struct ChunkRows : RowSource {
  int64_t C;
  const std::vector<int64_t> &reorder, &clen, &lens;
  const void *const *ccol;
  const void *const *cval;
  int index_type, in_vt;
  ChunkRows(int64_t c, const std::vector<int64_t> &r, const std::vector<int64_t> &cl, const std::vector<int64_t> &ln,
            const void *const *cc, const void *const *cv, int it, int iv)
      : C(c), reorder(r), clen(cl), lens(ln), ccol(cc), cval(cv), index_type(it), in_vt(iv) {}
  int64_t length(int64_t g) const override { return lens[g]; }
  int64_t orig_row(int64_t g) const override { return reorder[g]; }
  void entry(int64_t g, int64_t k, int64_t &col, double &val) const override {
    int64_t c = g / C, pos = (g % C) * clen[c] + k;
    col = index_type == SPOPS_INDEX_I32 ? static_cast<const int32_t *>(ccol[c])[pos]
                                        : static_cast<const int64_t *>(ccol[c])[pos];
    val = host_value(cval[c], in_vt, pos);
  }
};
}  // namespace

extern "C" int spops_sell_create_from_chunks(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int64_t C,
                                             int64_t nchunks, int index_type, const void *reorder,
                                             const void *chunk_len, const void *const *chunk_col,
                                             const void *const *chunk_val, int in_vt, int vt) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    SPOPS_REQUIRE(C > 0 && nchunks == (n_rows + C - 1) / C, SPOPS_ERROR_INVALID_VALUE,
                  "nchunks must be ceil(n_rows / C)");
    SPOPS_REQUIRE(n_rows < (int64_t)INT32_MAX && n_cols < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT,
                  "rows and columns must be < 2^31");
    std::vector<int64_t> perm = host_index(reorder, n_rows, index_type);
    std::vector<int64_t> clen = host_index(chunk_len, nchunks, index_type);
    std::vector<int64_t> lens(n_rows, 0);
    int64_t total = 0;
    for (int64_t g = 0; g < n_rows; ++g) {
      int64_t c = g / C, base = (g % C) * clen[c];
      int64_t k = 0;
      for (; k < clen[c]; ++k) {
        int64_t col = index_type == SPOPS_INDEX_I32 ? static_cast<const int32_t *>(chunk_col[c])[base + k]
                                                    : static_cast<const int64_t *>(chunk_col[c])[base + k];
        if (col < 0) break;
        SPOPS_REQUIRE(col < n_cols, SPOPS_ERROR_INVALID_VALUE, "column index out of range");
      }
      lens[g] = k;
      total += k;
      SPOPS_REQUIRE(perm[g] >= 0 && perm[g] < n_rows, SPOPS_ERROR_INVALID_VALUE, "reorder entry out of range");
    }
    SPOPS_REQUIRE(total == nnz, SPOPS_ERROR_INVALID_VALUE, "chunk entries (%lld) != nnz (%lld)", (long long)total,
                  (long long)nnz);
    auto M = std::make_unique<spops_matrix_s>();
    M->format = SPOPS_FORMAT_SELL;
    M->value_type = vt;
    M->index_type = index_type;
    M->n_rows = n_rows; M->n_cols = n_cols; M->nnz = nnz;
    ChunkRows src(C, perm, clen, lens, chunk_col, chunk_val, index_type, in_vt);
    build_sell_layout(*M, src, vt, device_chunk_height(C));
    M->sell.C = C;
    M->sell.sigma = 0;
    M->ref_nchunks = nchunks;
    M->sell.ref_reorder.assign(perm.begin(), perm.end());
    M->sell.ref_chunk_len.assign(clen.begin(), clen.end());
    M->algo_bytes = spmv_algo_bytes(*M, index_type == SPOPS_INDEX_I32 ? 4 : 8);
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_bsr_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                                const void *rowptr, const void *col, const void *values, int in_vt, int vt, int64_t bR,
                                int64_t bC) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    SPOPS_REQUIRE(bR > 0 && bC > 0, SPOPS_ERROR_INVALID_VALUE, "block dims must be positive");
    SPOPS_REQUIRE(bR * bC <= kBsrMaxBlock && bR <= kWave, SPOPS_ERROR_LIMIT, "block of %lld x %lld exceeds %d entries",
                  (long long)bR, (long long)bC, kBsrMaxBlock);
    std::vector<int64_t> rp = host_index(rowptr, n_rows + 1, index_type);
    std::vector<int64_t> ci = host_index(col, nnz, index_type);
    check_csr(n_rows, n_cols, nnz, rp, ci);
    const int64_t mb = (n_rows + bR - 1) / bR, nb = (n_cols + bC - 1) / bC;
    // csr_to_bsr (include/sparse_conversion.h:777-934): distinct block columns per block
    // row, ascending; block data row-major.
    std::vector<int32_t> brp(mb + 1, 0), bcol;
    std::vector<int64_t> mark(nb, -1);
    for (int64_t i = 0; i < mb; ++i) {
      int64_t hi = std::min(n_rows, (i + 1) * bR);
      std::vector<int64_t> cols;
      for (int64_t j = rp[i * bR]; j < rp[hi]; ++j) {
        int64_t b = ci[j] / bC;
        if (mark[b] != i) { mark[b] = i; cols.push_back(b); }
      }
      std::sort(cols.begin(), cols.end());
      for (int64_t b : cols) bcol.push_back((int32_t)b);
      brp[i + 1] = (int32_t)bcol.size();
    }
    const int64_t nnzb = (int64_t)bcol.size();
    SPOPS_REQUIRE(nnzb * bR * bC < ((int64_t)1 << 40), SPOPS_ERROR_LIMIT, "BSR too large");
    std::vector<double> dense((size_t)(nnzb * bR * bC), 0.0);
    for (int64_t r = 0; r < n_rows; ++r) {
      int64_t br = r / bR;
      for (int64_t j = rp[r]; j < rp[r + 1]; ++j) {
        int64_t b = ci[j] / bC;
        auto it = std::lower_bound(bcol.begin() + brp[br], bcol.begin() + brp[br + 1], (int32_t)b);
        int64_t k = it - bcol.begin();
        dense[(size_t)(k * bR * bC + (r % bR) * bC + ci[j] % bC)] = host_value(values, in_vt, j);
      }
    }
    std::vector<uint8_t> v;
    v.reserve(dense.size() * value_bytes(vt));
    for (double d : dense) pack_values(v, vt, d);
    auto M = std::make_unique<spops_matrix_s>();
    M->format = SPOPS_FORMAT_BSR;
    M->value_type = vt;
    M->index_type = index_type;
    M->n_rows = n_rows; M->n_cols = n_cols; M->nnz = nnz;
    M->bsr.brp.upload(brp);
    M->bsr.bcol.upload(bcol);
    M->bsr.data.upload(v);
    M->bsr.mb = mb; M->bsr.bR = bR; M->bsr.bC = bC; M->bsr.nnzb = nnzb;
    M->slots = nnzb * bR * bC;
    M->dev_bytes = M->bsr.brp.bytes + M->bsr.bcol.bytes + M->bsr.data.bytes;
    int64_t sV = (int64_t)value_bytes(vt), sX = (int64_t)xy_bytes(vt), sI = index_type == SPOPS_INDEX_I32 ? 4 : 8;
    M->algo_bytes = (mb + 1) * sI + nnzb * sI + nnzb * bR * bC * sV + n_cols * sX + n_rows * sX;
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmv(spops_matrix A, double alpha, const void *d_x, double beta, void *d_y, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null matrix");
    SPOPS_REQUIRE((d_x != nullptr || A->n_cols == 0) && (d_y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null x or y");
    launch_spmv(*A, alpha, d_x, beta, d_y, static_cast<hipStream_t>(stream));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_matrix_info(spops_matrix A, int64_t info[16]) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr && info != nullptr, SPOPS_ERROR_INVALID_VALUE, "null argument");
    for (int i = 0; i < 16; ++i) info[i] = 0;
    info[0] = A->format; info[1] = A->n_rows; info[2] = A->n_cols; info[3] = A->nnz; info[4] = A->value_type;
    info[5] = A->dev_bytes; info[6] = A->algo_bytes; info[7] = A->slots;
    info[15] = A->index_type;
    if (A->format == SPOPS_FORMAT_CSR) { info[8] = A->csr.nlong; info[9] = A->csr.ntasks; }
    if (A->format == SPOPS_FORMAT_SELL || A->format == SPOPS_FORMAT_ELL) {
      info[8] = A->sell.n_olong + A->sell.n_ovec; info[9] = A->sell.nchunks + A->sell.n_ovec;
      info[10] = A->sell.C; info[11] = A->sell.sigma; info[14] = A->ref_nchunks;
      if (A->format == SPOPS_FORMAT_ELL) info[13] = A->ell_width;
    }
    if (A->format == SPOPS_FORMAT_BSR) { info[12] = A->bsr.bR; info[13] = A->bsr.bC; info[14] = A->bsr.nnzb; }
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_matrix_layout(spops_matrix A, int which, void *host_out) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr && host_out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(A->format == SPOPS_FORMAT_SELL, SPOPS_ERROR_NOT_SUPPORTED, "layout arrays exist for SELL only");
    if (which == 0) std::memcpy(host_out, A->sell.ref_reorder.data(), A->sell.ref_reorder.size() * 4);
    else if (which == 1) std::memcpy(host_out, A->sell.ref_chunk_len.data(), A->sell.ref_chunk_len.size() * 4);
    else throw_error(SPOPS_ERROR_INVALID_VALUE, "unknown layout array %d", which);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_matrix_destroy(spops_matrix A) {
  delete A;
  return SPOPS_SUCCESS;
}

extern "C" int spops_spmm(spops_matrix A, int64_t k, double alpha, const void *d_X, double beta, void *d_Y,
                          void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr && k > 0, SPOPS_ERROR_INVALID_VALUE, "null matrix or k <= 0");
    (void)alpha; (void)d_X; (void)beta; (void)d_Y; (void)stream;
    throw_error(SPOPS_ERROR_NOT_SUPPORTED, "spops_spmm: not implemented for this format yet");
    return SPOPS_ERROR_NOT_SUPPORTED;
  });
}

extern "C" int spops_bsr_create_from_blocks(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t bR, int64_t bC,
                                            int64_t mb, int64_t nnzb, int index_type, const void *row_ptr,
                                            const void *block_colindex, const void *block_data, int in_vt, int vt) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    SPOPS_REQUIRE(bR > 0 && bC > 0 && mb == (n_rows + bR - 1) / bR, SPOPS_ERROR_INVALID_VALUE,
                  "mb must be ceil(n_rows / bR)");
    SPOPS_REQUIRE(bR * bC <= kBsrMaxBlock && bR <= kWave, SPOPS_ERROR_LIMIT, "block of %lld x %lld exceeds %d entries",
                  (long long)bR, (long long)bC, kBsrMaxBlock);
    SPOPS_REQUIRE((uint64_t)n_cols * 8 <= kMaxBufBytes, SPOPS_ERROR_LIMIT, "x must fit one 2 GiB buffer resource");
    std::vector<int64_t> rp = host_index(row_ptr, mb + 1, index_type);
    std::vector<int64_t> bc = host_index(block_colindex, nnzb, index_type);
    SPOPS_REQUIRE(rp[0] == 0 && rp[mb] == nnzb, SPOPS_ERROR_INVALID_VALUE, "row_ptr must run from 0 to nnzb");
    const int64_t nb = (n_cols + bC - 1) / bC;
    for (int64_t b : bc) SPOPS_REQUIRE(b >= 0 && b < nb, SPOPS_ERROR_INVALID_VALUE, "block column out of range");
    std::vector<uint8_t> v;
    v.reserve((size_t)(nnzb * bR * bC) * value_bytes(vt));
    int64_t nnz = 0;
    for (int64_t e = 0; e < nnzb * bR * bC; ++e) {
      double d = host_value(block_data, in_vt, e);
      nnz += d != 0.0;
      pack_values(v, vt, d);
    }
    auto M = std::make_unique<spops_matrix_s>();
    M->format = SPOPS_FORMAT_BSR;
    M->value_type = vt;
    M->index_type = index_type;
    M->n_rows = n_rows; M->n_cols = n_cols; M->nnz = nnz;
    M->bsr.brp.upload(std::vector<int32_t>(rp.begin(), rp.end()));
    M->bsr.bcol.upload(std::vector<int32_t>(bc.begin(), bc.end()));
    M->bsr.data.upload(v);
    M->bsr.mb = mb; M->bsr.bR = bR; M->bsr.bC = bC; M->bsr.nnzb = nnzb;
    M->slots = nnzb * bR * bC;
    M->dev_bytes = M->bsr.brp.bytes + M->bsr.bcol.bytes + M->bsr.data.bytes;
    int64_t sV = (int64_t)value_bytes(vt), sX = (int64_t)xy_bytes(vt), sI = index_type == SPOPS_INDEX_I32 ? 4 : 8;
    M->algo_bytes = (mb + 1) * sI + nnzb * sI + nnzb * bR * bC * sV + n_cols * sX + n_rows * sX;
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}
