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
//   SELL : (SELL-C-sigma and SELL-C-R) x is cut into P column slices of about equal entry
//          count, P the smallest power of two <= 8 that keeps a slice of x within 2 MiB
//          (one XCD's share of L2).  Workgroup g runs on XCD g % 8 and only tasks of slice
//          (g % 8) % P run there, so every XCD gathers from an L2-resident slice of x:
//          random gathers that hit L2 run at about twice the rate of gathers that miss
//          to the memory side.  Per slice the row segments are stored packed in row order
//          (no padding) with uint16 lengths; a wave task covers up to 64 rows and 512
//          entries (one lane per row), a segment longer than 64 entries is a wave task of
//          its own, rows longer than 2048 entries are workgroup tasks over all of x.
//          Products and sums are in double; each slice writes one double partial per
//          row and a second pass forms y = alpha * sum of partials + beta * y (P == 1:
//          the tasks write y directly).  The reference's chunk_len / reorder arrays are
//          kept on the host for inspection.
//   ELL  : the same layout.
//   SELL/ELL swept (matrices whose rows gather from all over x): blocks of consecutive rows
//          (<= 20480, equal entry counts) whose FP64 sums live in LDS; a block's entries
//          sorted by column, so the lanes of one gather read consecutive entries of one
//          sweep over x; H workgroups per block, each writing a partial plane that a second
//          pass adds in order (spmv_internal.hpp, SweepDev).
//   BSR  : the reference layout (row_ptr, block_colindex, row-major bR x bC blocks).
// The streams of A (indices, values) use non-temporal loads so that L2 keeps x.
#include <algorithm>
#include <type_traits>
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
template <int VT, typename A = typename VT_<VT>::Acc, bool kC = VT_<VT>::kComp>
__device__ __forceinline__ void row_sum_wave(const int *col, const typename VT_<VT>::V *val, int64_t len, rsrc_t rx,
                                             Accum<A, kC> &a) {
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
      for (int u = 0; u < 4; ++u) a.fma((A)T::cvt(v[u]), (A)xv[u]);
    }
  }
}

// Whole-workgroup sum of one long row; result valid in thread 0.
template <int VT, typename A = typename VT_<VT>::Acc, bool kC = VT_<VT>::kComp>
__device__ __forceinline__ A row_sum_block(const int *col, const typename VT_<VT>::V *val, int64_t len, rsrc_t rx,
                                           A *scratch /* 8 entries of LDS */) {
  using T = VT_<VT>;
  constexpr uint32_t sV = sizeof(typename T::V), sX = sizeof(typename T::X);
  Accum<A, kC> a;
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
      for (int u = 0; u < 4; ++u) a.fma((A)T::cvt(v[u]), (A)xv[u]);
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) a.merge_xor(m);
  if (lane_id() == 0) {
    scratch[wave_id()] = a.s;
    if constexpr (kC) scratch[4 + wave_id()] = a.c;
  }
  __syncthreads();
  A r = 0;
  if (threadIdx.x == 0) {
    Accum<A, kC> t;
    for (int w = 0; w < kBlock / kWave; ++w) {
      t.add(scratch[w]);
      if constexpr (kC) t.add(scratch[4 + w]);
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
// Column-sliced stream SpMV (SELL-C-sigma, SELL-C-R, ELL)
// grid = nlong_pad workgroups (one long row each, nlong_pad a multiple of 8) + 8 * wg_per_xcd
// workgroups of 4 wave tasks; workgroup nlong_pad + 8j + q takes tasks 4j.. of XCD queue q.
// ---------------------------------------------------------------------------------------------
struct SliceArgs {
  const int *col;
  const void *val;
  const uint32_t *meta;
  const int4 *tasks;
  const long *lptr;
  const int *lcol;
  const void *lval;
  const int *lrow;
  double *part;
  const void *x;
  void *y;
  long sbase[kXcds];
  int qoff[kXcds + 1];
  int n_rows, n_cols, P, nlong, nlong_pad;
  double alpha, beta;
};

template <typename X>
__device__ __forceinline__ void slice_store(const SliceArgs &a, bool direct, int s, int row, double sum) {
  if (direct) {
    X *y = static_cast<X *>(a.y);
    double r = a.alpha * sum;
    if (a.beta != 0.0) r += a.beta * (double)y[row];
    y[row] = (X)r;
  } else {
    a.part[(long)s * a.n_rows + row] = sum;
  }
}

// The K elements a lane owns in a task stored lane-interleaved: with g = 16 / sizeof(E)
// elements per 128-bit word, word b of lane l sits at word index b * L + l (L = lanes in
// the task), so each of the K / g loads of the wave reads one contiguous block.
template <typename E, int K>
__device__ __forceinline__ void load_lane_run(rsrc_t r, int L, int l, E (&out)[K]) {
  constexpr int NB = K * (int)sizeof(E) / 16;
  static_assert(NB * 16 == K * (int)sizeof(E), "run must be a multiple of 16 bytes");
  v4i raw[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) raw[b] = bload<v4i, kAuxNT>(r, l < L ? 16u * (uint32_t)(b * L + l) : kOOB);
  __builtin_memcpy(out, raw, sizeof(out));
}

// DPP lane moves (bound_ctrl: lanes without a source read 0; rows outside row_mask keep 0)
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, true);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = dpp<CTRL, ROWS>((int)(uint32_t)b), hi = dpp<CTRL, ROWS>((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

// One step of an inclusive segmented scan towards higher lanes: a lane whose range has no
// segment start yet adds the value arriving from below and takes over its start flag.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ void seg_scan_step(double &w, int &stop) {
  const double w2 = dpp<CTRL, ROWS>(w);
  const int f2 = dpp<CTRL, ROWS>(stop);
  if (!stop) {
    w += w2;
    stop = f2;
  }
}

template <int VT, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_slice_spmv(const SliceArgs a) {
  using T = VT_<VT>;
  using V = typename T::V;
  using X = typename T::X;
  constexpr uint32_t sV = sizeof(V), sX = sizeof(X);
  constexpr int K = slice_lane_entries(VT);
  __shared__ double red[8];
  __shared__ double rsum[kBlock / kWave][kSliceTaskRows];
  const rsrc_t rx = make_rsrc(a.x, (uint64_t)a.n_cols * sX);
  const int g = blockIdx.x;
  if (g < a.nlong_pad) {
    if (g >= a.nlong) return;
    const long p0 = a.lptr[g], p1 = a.lptr[g + 1];
    const double r = row_sum_block<VT, double, false>(a.lcol + p0, static_cast<const V *>(a.lval) + p0, p1 - p0,
                                                      rx, red);
    const int row = a.lrow[g];
    if (threadIdx.x == 0) slice_store<X>(a, DIRECT, 0, row, r);
    if (!DIRECT)
      for (int s = 1 + (int)threadIdx.x; s < a.P; s += kBlock) a.part[(long)s * a.n_rows + row] = 0.0;
    return;
  }
  const int gg = g - a.nlong_pad;
  const int q = gg % kXcds;
  const int t = __builtin_amdgcn_readfirstlane(a.qoff[q] + (gg / kXcds) * (kBlock / kWave) + wave_id());
  if (t >= a.qoff[q + 1]) return;
  // the task is wave-uniform: keep it in scalar registers so that every buffer resource
  // built from it is scalar too
  const int4 tk = a.tasks[t];
  const int r0 = __builtin_amdgcn_readfirstlane(tk.x), ne = __builtin_amdgcn_readfirstlane(tk.z);
  const int tw = __builtin_amdgcn_readfirstlane(tk.w);
  const int R = tw & 0xffff, s = (tw >> 16) & 0xff, kind = tw >> 24;
  const long eb = a.sbase[s] + __builtin_amdgcn_readfirstlane(tk.y);
  const int *col = a.col + eb;
  const V *val = static_cast<const V *>(a.val) + eb;
  const int lane = lane_id();
  if (kind == 1) {                          // one long segment: the wave strides over it
    Accum<double, false> acc;
    row_sum_wave<VT, double, false>(col, val, ne, rx, acc);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) acc.merge_xor(m);
    if (lane == 0) slice_store<X>(a, DIRECT, s, r0, acc.value());
    return;
  }
  // lane l: entries K l .. K l + K - 1 (dummy entries have column -1: their gather is out
  // of range and returns 0; lanes past the task load zeros and gather x[0], times 0)
  // Lane p holds the task's entry run ll = 63 - p, so that the pieces of a row spanning
  // several runs flow towards higher lanes (the direction of the DPP scan below).
  const int ll = kWave - 1 - lane;
  const uint32_t mw = a.meta[(long)t * kWave + ll];
  const int L = (ne + K - 1) / K;            // runs holding entries; the range is L * K long
  const rsrc_t rc = make_rsrc(col, (uint64_t)L * K * 4), rv = make_rsrc(val, (uint64_t)L * K * sV);
  int c[K];
  V v[K];
  load_lane_run<int, K>(rc, L, ll, c);
  load_lane_run<V, K>(rv, L, ll, v);
  X xv[K];
#pragma unroll
  for (int i = 0; i < K; ++i) xv[i] = bload<X>(rx, (uint32_t)c[i] * sX);
  // local pass: pieces that start and end inside the lane go to the task's row sums in LDS
  // at once; the piece before the lane's first row start (carry) and the one after its
  // last (tail) wait for the scan.
  double *rs = rsum[wave_id()];
  const uint32_t hbits = mw & 0xffffu;
  int row = (int)(mw >> 16) - 1;             // task-relative
  double cur = 0.0, carry = 0.0;
  bool seen = false;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (hbits & (1u << i)) {
      if (seen) rs[row] = cur;
      else carry = cur;
      cur = 0.0;
      ++row;
      seen = true;
    }
    cur = __builtin_fma((double)T::cvt(v[i]), (double)xv[i], cur);
  }
  if (!seen) carry = cur;
  // W_p = carry_p + (a row starts in lane p ? 0 : W_{p-1}); the row whose last start is in
  // lane p ends with cur_p + W_{p-1}
  double W = carry;
  int stop = seen ? 1 : 0;
  seg_scan_step<kDppRowShr + 1>(W, stop);
  seg_scan_step<kDppRowShr + 2>(W, stop);
  seg_scan_step<kDppRowShr + 4>(W, stop);
  seg_scan_step<kDppRowShr + 8>(W, stop);
  seg_scan_step<kDppRowBcast15, 0xa>(W, stop);
  seg_scan_step<kDppRowBcast31, 0xc>(W, stop);
  const double prev = dpp<kDppWaveShr1>(W);
  if (seen) rs[row] = cur + prev;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int j = lane; j < R; j += kWave) slice_store<X>(a, DIRECT, s, r0 + j, rs[j]);
}

// y = alpha * (sum over slices of part) + beta * y, slices added in order 0..P-1.
template <typename X>
__global__ __launch_bounds__(kBlock) void k_slice_reduce(const double *__restrict__ part, int P, int n_rows,
                                                        X *__restrict__ y, double alpha, double beta) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_rows) return;
  double sum = 0.0;
  for (int s = 0; s < P; ++s) sum += __builtin_nontemporal_load(part + (long)s * n_rows + r);
  double o = alpha * sum;
  if (beta != 0.0) o += beta * (double)y[r];
  y[r] = (X)o;
}

// ---------------------------------------------------------------------------------------------
// Swept layout: one workgroup of kSweepBlock threads per (row block, share h of its groups).
// ---------------------------------------------------------------------------------------------
struct SweepArgs {
  const uint32_t *word;
  const void *val;
  const int *gdesc, *wgrp;
  const int2 *wblk;           // per workgroup: {row block, share}
  const int4 *binfo;          // per row block: {first position, rows, shares, first plane element}
  const int *rowidx;          // position -> matrix row (nullptr: the identity)
  void *part;                 // blocks of > 1 share: y-type planes, plane h at first + h * stride
  const void *x;
  void *y;
  int n_rows, n_cols, rb, pstride;   // pstride 0: a block's planes are `rows` apart
  double alpha, beta;
};

template <int VT>
__global__ __launch_bounds__(kSweepBlock) void k_sweep_spmv(const SweepArgs a) {
  using T = VT_<VT>;
  using V = typename T::V;
  using X = typename T::X;
  constexpr uint32_t sV = sizeof(V), sX = sizeof(X);
  constexpr int K = kSweepLaneEntries, NW = kSweepBlock / kWave;
  // groups in flight per wave: two (four with the prefetch) when the values take one VGPR
  constexpr int D = sV == 8 ? 1 : 2;
  __shared__ double acc[kSweepMaxRows];
  const int rb = a.rb;
  const uint32_t rmask = (1u << rb) - 1u;
  const int2 wb = a.wblk[blockIdx.x];
  const int4 bi = a.binfo[wb.x];
  const int r0 = bi.x, rows = bi.y, h = wb.y;
  for (int i = (int)threadIdx.x; i < rows; i += kSweepBlock) acc[i] = 0.0;
  __syncthreads();
  const int g0 = a.wgrp[blockIdx.x], g1 = a.wgrp[blockIdx.x + 1];
  const int lane = lane_id();
  const long cb_cols = 1l << (32 - rb);
  // The 16 waves take D * 16 consecutive groups at a time (wave w: groups base + d * 16 + w),
  // so the workgroup sweeps x in order.  Each wave loads the words and values of its next D
  // groups after issuing the gathers of the current ones (loads return in issue order, so
  // the gathers are consumed while the stream loads are still in flight).  Groups past g1
  // load nothing and are skipped.  Padding slots (past `valid`) hold word 0 and value 0 and
  // gather nothing (out-of-range offset: 0), so they add an exact +0.0 to acc[0], which never
  // changes a sum that starts at +0.0.
  auto load_group = [&](int g, uint32_t (&w)[K], V (&v)[K]) {
    const bool in = g < g1;
    const rsrc_t rw = make_rsrc(a.word + (long)g * kSweepGroup, in ? (uint64_t)kSweepGroup * 4 : 0);
    const rsrc_t rv = make_rsrc(static_cast<const V *>(a.val) + (long)g * kSweepGroup, in ? (uint64_t)kSweepGroup * sV : 0);
    load_lane_run<uint32_t, K>(rw, kWave, lane, w);
    load_lane_run<V, K>(rv, kWave, lane, v);
    return in ? __builtin_amdgcn_readfirstlane(a.gdesc[g]) : 0;
  };
  int base = __builtin_amdgcn_readfirstlane(g0);
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  uint32_t w[D][K];
  V v[D][K];
  int d[D];
#pragma unroll
  for (int q = 0; q < D; ++q) d[q] = load_group(base + q * NW + wv, w[q], v[q]);
  for (; base < g1; base += D * NW) {
    X xv[D][K];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int cb = d[q] & 0xffff, valid = d[q] >> 16;
      const long c0 = (long)cb * cb_cols;
      const long ncx = (long)a.n_cols - c0 < cb_cols ? (long)a.n_cols - c0 : cb_cols;
      const rsrc_t rx = make_rsrc(static_cast<const X *>(a.x) + c0, (uint64_t)(ncx > 0 ? ncx : 0) * sX);
#pragma unroll
      for (int u = 0; u < K; ++u) xv[q][u] = bload<X>(rx, (u * kWave + lane < valid) ? (w[q][u] >> rb) * sX : kOOB);
    }
    uint32_t wn[D][K];
    V vn[D][K];
    int dn[D];
#pragma unroll
    for (int q = 0; q < D; ++q) dn[q] = load_group(base + (D + q) * NW + wv, wn[q], vn[q]);
#pragma unroll
    for (int q = 0; q < D; ++q)
      if (d[q] != 0) {
#pragma unroll
        for (int u = 0; u < K; ++u) atomicAdd(&acc[w[q][u] & rmask], (double)T::cvt(v[q][u]) * (double)xv[q][u]);
      }
#pragma unroll
    for (int q = 0; q < D; ++q) {
#pragma unroll
      for (int u = 0; u < K; ++u) { w[q][u] = wn[q][u]; v[q][u] = vn[q][u]; }
      d[q] = dn[q];
    }
  }
  __syncthreads();
  if (bi.z == 1) {
    X *y = static_cast<X *>(a.y);
    for (int i = (int)threadIdx.x; i < rows; i += kSweepBlock) {
      const int r = a.rowidx ? a.rowidx[r0 + i] : r0 + i;
      double o = a.alpha * acc[i];
      if (a.beta != 0.0) o += a.beta * (double)y[r];
      y[r] = (X)o;
    }
  } else {
    X *mine = static_cast<X *>(a.part) + bi.w + (long)h * (a.pstride ? a.pstride : rows);
    for (int i = (int)threadIdx.x; i < rows; i += kSweepBlock) mine[i] = (X)acc[i];
  }
}

// y = alpha * (sum of the H partial planes, in order 0..H-1, in double) + beta * y.  FP32
// planes round each block share's FP64 sum once: with H shares the error stays within
// (H + 1) / 2 units of FP32 rounding of the sum of |products|.
template <typename X>
__global__ __launch_bounds__(kBlock) void k_sweep_reduce(const X *__restrict__ part, int H, int n_rows,
                                                        X *__restrict__ y, double alpha, double beta) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_rows) return;
  double sum = 0.0;
  for (int h = 0; h < H; ++h) sum += (double)__builtin_nontemporal_load(part + (long)h * n_rows + r);
  double o = alpha * sum;
  if (beta != 0.0) o += beta * (double)y[r];
  y[r] = (X)o;
}

// The same for FP32 with four rows per thread (16-byte loads and stores; n_rows % 4 == 0 and
// 16-byte aligned planes and y)
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void k_sweep_reduce4(const f32x4 *__restrict__ part, int H, int n4,
                                                         f32x4 *__restrict__ y, double alpha, double beta) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n4) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int h = 0; h < H; ++h) {
    const f32x4 p = __builtin_nontemporal_load(part + (long)h * n4 + r);
    s0 += (double)p.x;
    s1 += (double)p.y;
    s2 += (double)p.z;
    s3 += (double)p.w;
  }
  double o0 = alpha * s0, o1 = alpha * s1, o2 = alpha * s2, o3 = alpha * s3;
  if (beta != 0.0) {
    const f32x4 old = y[r];
    o0 += beta * (double)old.x;
    o1 += beta * (double)old.y;
    o2 += beta * (double)old.z;
    o3 += beta * (double)old.w;
  }
  f32x4 o;
  o.x = (float)o0;
  o.y = (float)o1;
  o.z = (float)o2;
  o.w = (float)o3;
  y[r] = o;
}

// The same for a layout whose blocks have their own share counts and planes: one thread per
// row of a block of more than one share (task: {block, first local row, rows}); the row's
// planes are `rows` apart, added in order h = 0..shares-1 in double.
template <typename X>
__global__ __launch_bounds__(kBlock) void k_sweep_gather(const int4 *__restrict__ tasks, const int4 *__restrict__ binfo,
                                                        const X *__restrict__ part, const int *__restrict__ rowidx,
                                                        X *__restrict__ y, double alpha, double beta) {
  const int4 t = tasks[blockIdx.x];
  if ((int)threadIdx.x >= t.z) return;
  const int4 bi = binfo[t.x];
  const int local = t.y + (int)threadIdx.x;
  const X *p = part + bi.w + local;
  double sum = 0.0;
  for (int h = 0; h < bi.z; ++h) sum += (double)__builtin_nontemporal_load(p + (long)h * bi.y);
  const int r = rowidx ? rowidx[bi.x + local] : bi.x + local;
  double o = alpha * sum;
  if (beta != 0.0) o += beta * (double)y[r];
  y[r] = (X)o;
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

// BSR with dense blocks of at least 8 x 8 (bC % 4 == 0), on the matrix cores: one wave per
// block row; each block is cut into tiles of 16 rows x 16 columns, each tile four
// v_mfma_f64_16x16x4f64 steps.  The reduction index is permuted alike in both operands: at
// step s, lane i + 16k gives A[i][4k + s] (so lane i + 16k loads the four consecutive
// entries 4k .. 4k+3 of row i of the tile in one vector load) and x[4k + s] (the same in
// all 16 columns of the B operand, so every column of D holds the tile's y).  Products are
// exact in FP64 for FP32 and BF16 values, sums FP64.  Rows past bR, columns past bC and x
// past n_cols read 0 (out-of-range buffer offsets).  D of this FP64 instruction: lane l,
// register r holds row l / 16 + 4 r of column l % 16 (measured on gfx950).
typedef double d4_t __attribute__((ext_vector_type(4)));

// MT: 16-row tiles per block (bR <= 16 MT); two accumulator chains per tile (steps s = 0, 2
// and s = 1, 3) so that consecutive MFMAs do not wait on each other.
template <int VT, int MT>
__global__ __launch_bounds__(kBlock) void k_bsr_mfma(const int *__restrict__ brp, const int *__restrict__ bcol,
                                                     const typename VT_<VT>::V *__restrict__ bdata, int mb, int bR,
                                                     int bC, int n_rows, int n_cols,
                                                     const typename VT_<VT>::X *__restrict__ x,
                                                     typename VT_<VT>::X *__restrict__ y, double alpha, double beta) {
  using T = VT_<VT>;
  using V = typename T::V;
  using X = typename T::X;
  constexpr uint32_t sV = sizeof(V), sX = sizeof(X);
  const int i = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (kBlock / kWave) + wave_id());
  if (i >= mb) return;
  const int lane = lane_id(), li = lane & 15, lk = lane >> 4;
  const int tiles = (bR + 15) >> 4, groups = (bC + 15) >> 4;
  const rsrc_t rx = make_rsrc(x, (uint64_t)n_cols * sX);
  const int b0 = brp[i], b1 = brp[i + 1];
  const int bn = bR * bC;
  const rsrc_t rd = make_rsrc(bdata + (long)b0 * bn, (uint64_t)(b1 - b0) * bn * sV);
  d4_t acc[MT][2];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t][0] = acc[t][1] = d4_t{0.0, 0.0, 0.0, 0.0};
  for (int b = b0; b < b1; ++b) {
    const long xb = (long)bcol[b] * bC;
    for (int cg = 0; cg < groups; ++cg) {
      const int c = cg * 16 + 4 * lk;              // this lane's first column of the group
      double xv[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const long xi = xb + c + s;
        xv[s] = (double)bload<X>(rx, (c < bC && xi < n_cols) ? (uint32_t)xi * sX : kOOB);
      }
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        if (t >= tiles) break;
        const int row = t * 16 + li;
        const uint32_t off = (row < bR && c < bC) ? (uint32_t)((b - b0) * bn + row * bC + c) * sV : kOOB;
        double av[4];
        if constexpr (sV == 8) {
          const bu4 lo = bload<bu4, kAuxNT>(rd, off), hi = bload<bu4, kAuxNT>(rd, off == kOOB ? kOOB : off + 16);
          __builtin_memcpy(&av[0], &lo, 16);
          __builtin_memcpy(&av[2], &hi, 16);
        } else if constexpr (sV == 4) {
          const bu4 w = bload<bu4, kAuxNT>(rd, off);
          float f[4];
          __builtin_memcpy(f, &w, 16);
#pragma unroll
          for (int s = 0; s < 4; ++s) av[s] = (double)f[s];
        } else {
          const bu2 w = bload<bu2, kAuxNT>(rd, off);
          uint16_t h[4];
          __builtin_memcpy(h, &w, 8);
#pragma unroll
          for (int s = 0; s < 4; ++s) av[s] = (double)T::cvt(h[s]);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
          acc[t][s & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], xv[s], acc[t][s & 1], 0, 0, 0);
      }
    }
  }
  if (li != 0) return;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    if (t >= tiles) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = t * 16 + lk + 4 * r;
      const long gr = (long)i * bR + row;
      if (row < bR && gr < n_rows) {
        double o = alpha * (acc[t][0][r] + acc[t][1][r]);
        if (beta != 0.0) o += beta * (double)y[gr];
        y[gr] = (X)o;
      }
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

void upload_csr_plan(CsrDev &P, const std::vector<int64_t> &rp, int64_t n_rows) {
  std::vector<int2> tasks, segs, sdt;
  std::vector<int> longs, order, slongs, segoff{0}, win;
  build_csr_plan(rp, n_rows, tasks, longs);
  for (int64_t w0 = 0; w0 < n_rows; w0 += kSpmmWindow) {
    const size_t at = order.size();
    win.clear();
    for (int64_t r = w0; r < std::min(n_rows, w0 + kSpmmWindow); ++r) {
      const int64_t len = rp[r + 1] - rp[r];
      if (len > 0) win.push_back((int)r);
      if (len <= kSpmmLongRow) { order.push_back((int)r); continue; }
      slongs.push_back((int)r);
      for (int64_t off = 0; off < len; off += kSpmmSegment) segs.push_back(make_int2((int)slongs.size() - 1, (int)off));
      segoff.push_back((int)segs.size());
    }
    auto longer = [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; };
    std::stable_sort(order.begin() + at, order.end(), longer);
    // SDDMM: every row with entries in chunks of at most kSddmmChunk, longest rows first
    std::stable_sort(win.begin(), win.end(), longer);
    for (int r : win)
      for (int64_t off = 0; off < rp[r + 1] - rp[r]; off += kSddmmChunk) sdt.push_back(make_int2(r, (int)off));
  }
  P.sdtasks.upload(sdt);
  P.nsdtasks = (int64_t)sdt.size();
  P.tasks.upload(tasks);
  P.longs.upload(longs);
  P.sorder.upload(order);
  P.slongs.upload(slongs);
  P.lsegs.upload(segs);
  P.lsegoff.upload(segoff);
  P.ntasks = (int)tasks.size();
  P.nlong = (int)longs.size();
  P.nshort = (int)order.size();
  P.nslong = (int)slongs.size();
  P.nlseg = (int)segs.size();
}

// Row access for the sliced builder: entries of the row at SELL position g.
struct RowSource {
  virtual ~RowSource() = default;
  virtual int64_t length(int64_t g) const = 0;
  virtual int64_t orig_row(int64_t g) const = 0;
  virtual void entry(int64_t g, int64_t k, int64_t &col, double &val) const = 0;
  // index of the entry in the source arrays (CSR sources; -1 when there is none)
  virtual int64_t entry_index(int64_t g, int64_t k) const { (void)g; (void)k; return -1; }
};

int64_t g_slice_bytes = kSliceBytesDefault;
int64_t g_dense_products = (int64_t)1 << 22;
int64_t g_sweep_mode = 0;
int64_t g_sweep_workgroups = kSweepTargetWorkgroups;
int64_t g_sweep_max_rows = kSweepMaxRows;
int64_t g_sweep_split = 0;
int64_t g_bsr_mfma = 1;   // 0 never, 1 blocks past the VALU kernel's limit, 2 every dense block >= 8 x 8

bool bsr_use_mfma(int64_t bR, int64_t bC) {
  const bool dense = bR >= kBsrMfmaMinDim && bC >= kBsrMfmaMinDim && bC % 4 == 0 && bR <= kBsrMfmaMaxRows &&
                     bR * bC <= kBsrMfmaMaxBlock;
  const bool valu = bR * bC <= kBsrMaxBlock && bR <= kWave;
  // one right-hand side leaves the matrix cores 1/16 of each 16x16x4 step: on 16x16 blocks
  // the VALU kernel is faster (FP64 56.9 vs 62.4 us, FP32 26.0 vs 38.5 us on 33.5 M entries),
  // so by default the MFMA kernel takes the dense blocks the VALU kernel cannot hold
  if (dense && (g_bsr_mfma == 2 || (g_bsr_mfma == 1 && !valu))) return true;
  SPOPS_REQUIRE(bR * bC <= kBsrMaxBlock && bR <= kWave, SPOPS_ERROR_LIMIT,
                "block of %lld x %lld: at most %d entries (or 8..%d x 8.. with bC %% 4 == 0 and at most %d entries "
                "for the MFMA kernel)", (long long)bR, (long long)bC, kBsrMaxBlock, kBsrMfmaMaxRows, kBsrMfmaMaxBlock);
  return false;
}

// Stable LSD radix sort of idx[0..n) by key[idx] (keys < 2^bits), 11-bit digits.
static void radix_sort_by(std::vector<int32_t> &idx, const std::vector<uint32_t> &key, int bits) {
  const size_t n = idx.size();
  std::vector<int32_t> tmp(n);
  std::vector<size_t> cnt(2048 + 1);
  for (int sh = 0; sh < bits; sh += 11) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) ++cnt[((key[idx[i]] >> sh) & 2047u) + 1];
    for (int d = 0; d < 2048; ++d) cnt[d + 1] += cnt[d];
    for (size_t i = 0; i < n; ++i) tmp[cnt[(key[idx[i]] >> sh) & 2047u]++] = idx[i];
    idx.swap(tmp);
  }
}

// The blocks of a swept layout: positions (row order of the blocks), block starts, shares.
struct SweepPlan {
  std::vector<int32_t> order;      // position -> row; empty: the identity
  std::vector<int32_t> rstart{0};  // per block + 1: first position
  std::vector<int> shares;         // per block
  int64_t split = 0;               // 0: one class
};

// One class: about btarget blocks of consecutive rows with equal entry counts (each at most
// rmax rows), H shares each.
static SweepPlan sweep_plan_one_class(const std::vector<int64_t> &cum, int64_t n, int64_t rmax, int64_t wgs) {
  SweepPlan P;
  const int64_t bmin = std::max<int64_t>(1, (n + rmax - 1) / rmax);
  const int H = (int)std::min<int64_t>(kSweepMaxSplit, std::max<int64_t>(1, wgs / bmin));
  const int64_t btarget = std::max<int64_t>(bmin, wgs / H);
  // greedy cut: a block takes rows while it has at most `cap` entries and rmax rows (and at
  // least one row); the smallest cap that needs at most btarget blocks, by bisection (the
  // greedy count falls as cap grows, and cap = nnz needs bmin <= btarget blocks)
  auto cut = [&](int64_t cap, std::vector<int32_t> *out) {
    int64_t blocks = 0;
    for (int64_t r = 0; r < n; ++blocks) {
      const int64_t limit = std::min(n, r + rmax);
      int64_t r1 = std::upper_bound(cum.begin() + r + 1, cum.begin() + limit + 1, cum[r] + cap) - cum.begin() - 1;
      r1 = std::max(r1, r + 1);
      if (out) out->push_back((int32_t)r1);
      r = r1;
    }
    return blocks;
  };
  int64_t lo = std::max<int64_t>(1, cum[n] / btarget), hi = std::max<int64_t>(1, cum[n]);
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cut(mid, nullptr) <= btarget) hi = mid;
    else lo = mid + 1;
  }
  cut(lo, &P.rstart);
  if (n == 0) P.rstart.push_back(0);
  P.shares.assign(P.rstart.size() - 1, H);
  return P;
}

// Two classes: rows of at most T entries, then the longer rows (each run in row order), cut
// into blocks of about rmax rows; shares handed out one at a time to the block with the
// largest load (entries per share + rows) until `wgs` workgroups are used.
static SweepPlan sweep_plan_two_class(const std::vector<int64_t> &len, int64_t n, int64_t rmax, int64_t wgs,
                                      int64_t T) {
  SweepPlan P;
  P.split = T;
  P.order.reserve((size_t)n);
  for (int pass = 0; pass < 2; ++pass)
    for (int64_t r = 0; r < n; ++r)
      if ((len[r] <= T) == (pass == 0)) P.order.push_back((int32_t)r);
  std::vector<int64_t> E;
  int64_t p0 = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int64_t nc = 0;
    for (int64_t r = 0; r < n; ++r) nc += (len[r] <= T) == (pass == 0);
    const int64_t nb = (nc + rmax - 1) / rmax;
    for (int64_t k = 0; k < nb; ++k) {
      const int64_t a = p0 + nc * k / nb, b = p0 + nc * (k + 1) / nb;
      int64_t e = 0;
      for (int64_t q = a; q < b; ++q) e += len[P.order[(size_t)q]];
      P.rstart.push_back((int32_t)b);
      E.push_back(e);
    }
    p0 += nc;
  }
  if (n == 0) { P.rstart.push_back(0); E.push_back(0); }
  const size_t nb = E.size();
  P.shares.assign(nb, 1);
  auto load = [&](size_t b) {
    return (double)E[b] / P.shares[b] + (double)(P.rstart[b + 1] - P.rstart[b]);
  };
  std::vector<std::pair<double, size_t>> heap;
  for (size_t b = 0; b < nb; ++b) heap.push_back({load(b), b});
  std::make_heap(heap.begin(), heap.end());
  for (int64_t used = (int64_t)nb; used < wgs && !heap.empty(); ++used) {
    std::pop_heap(heap.begin(), heap.end());
    const size_t b = heap.back().second;
    heap.pop_back();
    if (P.shares[b] >= kSweepMaxShares || P.shares[b] * (int64_t)kSweepGroup >= E[b]) { --used; continue; }
    ++P.shares[b];
    heap.push_back({load(b), b});
    std::push_heap(heap.begin(), heap.end());
  }
  return P;
}

// The swept layout (see SweepDev).  pos[r] is the SELL position of original row r, len[r] its length.
// With slot_src, the source entry of every slot (-1 for padding) is kept on the device, so
// that new values of the source arrays can be copied into the layout (CSR handles).
static void build_sweep_layout(spops_matrix_s &M, const RowSource &src, const std::vector<int64_t> &pos,
                               const std::vector<int64_t> &len, int vt, bool slot_src = false) {
  auto &W = M.sweep;
  const int64_t n = M.n_rows, m = M.n_cols;
  const int64_t rmax = g_sweep_max_rows;
  std::vector<int64_t> cum(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) cum[r + 1] = cum[r] + len[r];
  const SweepPlan plan = g_sweep_split > 0 ? sweep_plan_two_class(len, n, rmax, g_sweep_workgroups, g_sweep_split)
                                            : sweep_plan_one_class(cum, n, rmax, g_sweep_workgroups);
  const std::vector<int32_t> &rstart = plan.rstart;
  const std::vector<int32_t> &order = plan.order;
  const int64_t n_rb = (int64_t)rstart.size() - 1;
  int64_t maxrows = 1;
  for (int64_t b = 0; b < n_rb; ++b) maxrows = std::max<int64_t>(maxrows, rstart[b + 1] - rstart[b]);
  int rb = 6;
  while (((int64_t)1 << rb) < maxrows) ++rb;
  const int cbits = 32 - rb;
  SPOPS_REQUIRE((m >> cbits) < 0xffff, SPOPS_ERROR_LIMIT, "too many column blocks for the swept layout");
  int colbits = 1;
  while (colbits < 31 && ((int64_t)1 << colbits) < m) ++colbits;
  const int gc = 4, gv = 16 / (int)value_bytes(vt);
  std::vector<uint32_t> word;
  std::vector<double> dval;
  std::vector<int32_t> gdesc;
  std::vector<int64_t> bg(n_rb + 1, 0);                            // first group of each row block
  std::vector<int32_t> idx;
  std::vector<uint32_t> bcol, bword;
  std::vector<double> bval;
  std::vector<int32_t> bsrc, ssrc;
  for (int64_t b = 0; b < n_rb; ++b) {
    bcol.clear(); bword.clear(); bval.clear(); bsrc.clear();
    for (int64_t q = rstart[b]; q < rstart[b + 1]; ++q) {
      const int64_t r = order.empty() ? q : order[(size_t)q];
      for (int64_t k = 0; k < len[r]; ++k) {
        int64_t c; double v;
        src.entry(pos[r], k, c, v);
        bcol.push_back((uint32_t)c);
        bword.push_back((uint32_t)(((uint64_t)c & (((uint64_t)1 << cbits) - 1)) << rb) | (uint32_t)(q - rstart[b]));
        bval.push_back(v);
        if (slot_src) bsrc.push_back((int32_t)src.entry_index(pos[r], k));
      }
    }
    idx.resize(bcol.size());
    std::iota(idx.begin(), idx.end(), 0);
    radix_sort_by(idx, bcol, colbits);                             // by column; rows ascend among equals
    size_t i = 0;
    while (i < idx.size()) {
      const uint32_t cb = bcol[idx[i]] >> cbits;
      size_t e = i;
      while (e < idx.size() && (bcol[idx[e]] >> cbits) == cb) ++e;
      for (size_t g0 = i; g0 < e; g0 += kSweepGroup) {
        const int valid = (int)std::min<size_t>(kSweepGroup, e - g0);
        gdesc.push_back((int32_t)cb | (valid << 16));
        const size_t at = word.size();
        word.resize(at + kSweepGroup, 0u);
        dval.resize(at + kSweepGroup, 0.0);
        if (slot_src) ssrc.resize(at + kSweepGroup, -1);
        for (int j = 0; j < valid; ++j) {
          const int l = j % kWave, u = j / kWave;
          const int32_t s = idx[g0 + j];
          word[at + (size_t)((u / gc) * kWave + l) * gc + u % gc] = bword[s];
          dval[at + (size_t)((u / gv) * kWave + l) * gv + u % gv] = bval[s];
          if (slot_src) ssrc[at + (size_t)((u / gv) * kWave + l) * gv + u % gv] = bsrc[s];
        }
      }
      i = e;
    }
    bg[b + 1] = (int64_t)gdesc.size();
  }
  const int64_t ngroups = (int64_t)gdesc.size();
  SPOPS_REQUIRE(ngroups * kSweepGroup < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT,
                "the swept layout holds >= 2^31 entry slots");
  // workgroups (block-major, shares consecutive) and the planes of the blocks of several shares
  const bool one_class = order.empty();
  std::vector<int32_t> wgrp;
  std::vector<int2> wblk;
  std::vector<int4> binfo((size_t)n_rb), rtasks;
  int64_t pel = 0;
  int maxH = 1;
  for (int64_t b = 0; b < n_rb; ++b) {
    const int Hb = ngroups == 0 ? 1 : plan.shares[(size_t)b];
    const int32_t rows = rstart[b + 1] - rstart[b];
    maxH = std::max(maxH, Hb);
    for (int h = 0; h < Hb; ++h) {
      wgrp.push_back((int32_t)(bg[b] + (bg[b + 1] - bg[b]) * h / Hb));
      wblk.push_back(make_int2((int)b, h));
    }
    int64_t first = 0;
    if (Hb > 1) {
      if (one_class) {
        first = rstart[b];                                         // planes [H][n_rows]
      } else {
        first = pel;                                               // planes [shares][rows] per block
        pel += (int64_t)Hb * rows;
        for (int32_t l0 = 0; l0 < rows; l0 += kBlock) rtasks.push_back(make_int4((int)b, l0, std::min(kBlock, rows - l0), 0));
      }
    }
    binfo[(size_t)b] = make_int4(rstart[b], rows, Hb, (int)first);
  }
  if (one_class && maxH > 1) pel = (int64_t)maxH * n;
  SPOPS_REQUIRE(pel < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT, "the swept layout's partial planes hold >= 2^31 elements");
  wgrp.push_back((int32_t)ngroups);
  std::vector<uint8_t> val;
  val.reserve(dval.size() * value_bytes(vt));
  for (double v : dval) pack_values(val, vt, v);
  W.word.upload(word);
  W.val.upload(val);
  if (slot_src) W.src.upload(ssrc);
  W.gdesc.upload(gdesc);
  W.wgrp.upload(wgrp);
  W.wblk.upload(wblk);
  W.binfo.upload(binfo);
  if (!one_class) {
    W.rowidx.upload(order);
    W.rtasks.upload(rtasks);
  }
  if (pel > 0) W.part.alloc((size_t)pel * xy_bytes(vt));
  W.rb = rb;
  W.H = maxH;
  W.n_rb = (int)n_rb;
  W.ngroups = (int)ngroups;
  W.nwg = (int)wblk.size();
  W.nrtasks = (int)rtasks.size();
  W.split = plan.split;
  W.plane_elems = pel;
  M.slots = ngroups * kSweepGroup;
  M.dev_bytes = W.word.bytes + W.val.bytes + W.src.bytes + W.gdesc.bytes + W.wgrp.bytes + W.wblk.bytes +
                W.binfo.bytes + W.rowidx.bytes + W.rtasks.bytes + W.part.bytes;
}

// Column slices of the sliced layout: P from the x footprint of one XCD's rows (128-byte lines
// of x touched by an eighth of the rows, the largest eighth).
static int slice_count(const spops_matrix_s &M, const RowSource &src, const std::vector<int64_t> &pos,
                       const std::vector<int64_t> &len, int vt) {
  const int64_t n = M.n_rows, m = M.n_cols;
  const int64_t sx = (int64_t)xy_bytes(vt);
  int64_t foot = 0;
  if (m * sx > g_slice_bytes) {
    const int64_t per_line = 128 / sx, nlines = (m + per_line - 1) / per_line;
    std::vector<uint64_t> seen((size_t)(nlines + 63) / 64);
    int64_t tot = 0;
    for (int64_t r = 0; r < n; ++r) tot += len[r] > kCsrLongRow ? 0 : len[r];
    int64_t acc = 0, r = 0;
    for (int part = 0; part < kXcds; ++part) {
      std::fill(seen.begin(), seen.end(), 0ull);
      int64_t lines = 0;
      for (; r < n && acc * kXcds < (part + 1) * tot; ++r) {
        if (len[r] > kCsrLongRow) continue;
        for (int64_t k = 0; k < len[r]; ++k) {
          int64_t c; double v;
          src.entry(pos[r], k, c, v);
          const int64_t ln = c / per_line;
          const uint64_t bit = 1ull << (ln & 63);
          if (!(seen[(size_t)(ln >> 6)] & bit)) { seen[(size_t)(ln >> 6)] |= bit; ++lines; }
        }
        acc += len[r];
      }
      foot = std::max(foot, lines * 128);
    }
  }
  int P = 1;
  while (P < kXcds && foot > (int64_t)P * g_slice_bytes) P *= 2;
  return P;
}

// Rows that gather from more x than one XCD's L2 holds (P > 1) take the swept layout when a
// row block of kSweepRowsTypical rows holds at least one entry per 128-byte line of x (so the
// sorted sweep reads every line it fetches more than once on average); g_sweep_mode 1 / 2
// force the other layout / the swept one.
static bool sweep_wanted(const spops_matrix_s &M, int P, int vt) {
  const int64_t n = M.n_rows, m = M.n_cols, sx = (int64_t)xy_bytes(vt);
  return g_sweep_mode == 2 || (g_sweep_mode == 0 && P > 1 && n > 0 &&
                               (double)M.nnz * (double)std::min<int64_t>(n, kSweepRowsTypical) / (double)n >=
                                   (double)((m * sx + 127) / 128));
}

void build_slice_layout(spops_matrix_s &M, const RowSource &src, int vt) {
  auto &S = M.sell;
  const int64_t n = M.n_rows, m = M.n_cols;
  std::vector<int64_t> pos(n);                          // original row -> SELL position
  for (int64_t g = 0; g < n; ++g) pos[src.orig_row(g)] = g;
  std::vector<int64_t> len(n);
  for (int64_t r = 0; r < n; ++r) len[r] = src.length(pos[r]);
  const int P = slice_count(M, src, pos, len, vt);
  if (sweep_wanted(M, P, vt)) {
    build_sweep_layout(M, src, pos, len, vt);
    return;
  }
  std::vector<int64_t> cut(P + 1, m);
  cut[0] = 0;
  if (P > 1) {
    std::vector<int64_t> hist(m + 1, 0);
    int64_t tot = 0;
    for (int64_t r = 0; r < n; ++r) {
      if (len[r] > kCsrLongRow) continue;
      for (int64_t k = 0; k < len[r]; ++k) {
        int64_t c; double v;
        src.entry(pos[r], k, c, v);
        ++hist[c];
      }
      tot += len[r];
    }
    int64_t acc = 0, sl = 1;
    for (int64_t c = 0; c < m && sl < P; ++c) {
      acc += hist[c];
      while (sl < P && acc * P >= sl * tot) cut[sl++] = c + 1;
    }
    for (int i = 1; i <= P; ++i) cut[i] = std::max(cut[i], cut[i - 1]);
  }
  auto slice_of = [&](int64_t c) {
    return (int)(std::upper_bound(cut.begin() + 1, cut.end(), c) - (cut.begin() + 1));
  };
  // segments: entries of each row bucketed per slice (temporary, slice-major, row order)
  std::vector<int32_t> seg((size_t)P * n, 0);
  std::vector<int64_t> lrows;
  for (int64_t r = 0; r < n; ++r) {
    if (len[r] > kCsrLongRow) { lrows.push_back(r); continue; }
    for (int64_t k = 0; k < len[r]; ++k) {
      int64_t c; double v;
      src.entry(pos[r], k, c, v);
      ++seg[(size_t)slice_of(c) * n + r];
    }
  }
  std::vector<int64_t> seg_off((size_t)P * n + 1, 0);
  for (size_t i = 0; i < (size_t)P * n; ++i) seg_off[i + 1] = seg_off[i] + seg[i];
  const int64_t total = seg_off[(size_t)P * n];
  std::vector<int32_t> tcol((size_t)total);
  std::vector<double> tval((size_t)total);
  {
    std::vector<int64_t> fill(seg_off.begin(), seg_off.end() - 1);
    for (int64_t r = 0; r < n; ++r) {
      if (len[r] > kCsrLongRow) continue;
      for (int64_t k = 0; k < len[r]; ++k) {
        int64_t c; double v;
        src.entry(pos[r], k, c, v);
        const int64_t at = fill[(size_t)slice_of(c) * n + r]++;
        tcol[at] = (int32_t)c;
        tval[at] = v;
      }
    }
  }
  // tasks and the final entry order: every non-long row has a segment of >= 1 entry in every
  // slice (a dummy with column -1 when empty); each task's range is padded with dummies to a
  // multiple of K entries (so task ranges start 16-byte aligned) and lane-interleaved.
  const int K = slice_lane_entries(vt);
  std::vector<int32_t> col;
  std::vector<double> dval;
  col.reserve((size_t)(total + total / 8 + 64));
  dval.reserve(col.capacity());
  std::vector<std::vector<int4>> stasks(P);
  std::vector<std::vector<int64_t>> sents(P);
  std::vector<std::vector<uint32_t>> smeta(P);          // 64 words per task
  for (int s = 0; s < P; ++s) {
    S.sbase[s] = (int64_t)col.size();
    const int64_t s0 = S.sbase[s];
    int64_t r0 = 0, e0 = 0;
    int rows = 0, ne = 0;
    uint32_t lanemeta[kWave];
    auto pad = [&]() {
      while (col.size() % K) { col.push_back(-1); dval.push_back(0.0); }
    };
    // lane-interleave the task's entries (see load_lane_run): element i of lane l goes to
    // ((i / g) * L + l) * g + i % g, g = elements per 16 bytes (4 for columns, 16 / sV values)
    auto interleave = [&](int64_t at, int L) {
      const int gc = 4, gv = 16 / (int)value_bytes(vt);
      std::vector<int32_t> c2((size_t)L * K);
      std::vector<double> v2((size_t)L * K);
      for (int l = 0; l < L; ++l)
        for (int i = 0; i < K; ++i) {
          const size_t from = (size_t)at + (size_t)l * K + i;
          c2[(size_t)((i / gc) * L + l) * gc + i % gc] = col[from];
          v2[(size_t)((i / gv) * L + l) * gv + i % gv] = dval[from];
        }
      std::copy(c2.begin(), c2.end(), col.begin() + at);
      std::copy(v2.begin(), v2.end(), dval.begin() + at);
    };
    auto close = [&]() {
      if (rows > 0) {
        stasks[s].push_back(make_int4((int)r0, (int)(e0 - s0), ne, rows | (s << 16)));
        sents[s].push_back(ne);
        // lane words: row-start bits and rows started in earlier lanes
        uint32_t before = 0;
        for (int l = 0; l < kWave; ++l) {
          const uint32_t bits = lanemeta[l];
          smeta[s].push_back(bits | (before << 16));
          before += (uint32_t)__builtin_popcount(bits);
        }
        pad();
        interleave(e0, (ne + K - 1) / K);
      }
      rows = 0;
      ne = 0;
    };
    auto append = [&](int64_t r, int64_t l) {
      const int64_t so = seg_off[(size_t)s * n + r];
      col.insert(col.end(), tcol.begin() + so, tcol.begin() + so + l);
      dval.insert(dval.end(), tval.begin() + so, tval.begin() + so + l);
    };
    for (int64_t r = 0; r < n; ++r) {
      if (len[r] > kCsrLongRow) { close(); continue; }
      const int64_t l = seg[(size_t)s * n + r];
      if (l > kWave * K) {                              // a wave task of its own
        close();
        stasks[s].push_back(make_int4((int)r, (int)((int64_t)col.size() - s0), (int)l, (s << 16) | (1 << 24)));
        sents[s].push_back(l);
        smeta[s].insert(smeta[s].end(), kWave, 0u);
        append(r, l);
        pad();
        continue;
      }
      const int64_t le = l > 0 ? l : 1;
      if (ne + le > kWave * K || rows == kSliceTaskRows) close();
      if (rows == 0) {
        r0 = r;
        e0 = (int64_t)col.size();
        for (auto &w : lanemeta) w = 0;
      }
      lanemeta[ne / K] |= 1u << (ne % K);              // the row starts at entry ne
      ++rows;
      if (l > 0) append(r, l);
      else { col.push_back(-1); dval.push_back(0.0); }
      ne += (int)le;
    }
    close();
    SPOPS_REQUIRE((int64_t)col.size() - s0 < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT,
                  "a column slice holds >= 2^31 entries");
  }
  // XCD queues: queue q serves slice q % P, part q / P of its tasks (equal entry counts)
  std::vector<int4> tasks;
  std::vector<uint32_t> meta;
  const int parts = kXcds / P;
  for (int q = 0; q < kXcds; ++q) {
    S.qoff[q] = (int)tasks.size();
    const int s = q % P, k = q / P;
    int64_t tot = 0;
    for (int64_t e : sents[s]) tot += e + 16;           // + a fixed cost per task
    int64_t acc = 0;
    for (size_t i = 0; i < stasks[s].size(); ++i) {
      const int64_t part = std::min<int64_t>(parts - 1, acc * parts / std::max<int64_t>(tot, 1));
      acc += sents[s][i] + 16;
      if (part == k) {
        tasks.push_back(stasks[s][i]);
        meta.insert(meta.end(), smeta[s].begin() + (int64_t)i * kWave, smeta[s].begin() + (int64_t)(i + 1) * kWave);
      }
    }
  }
  S.qoff[kXcds] = (int)tasks.size();
  // long rows
  std::vector<long> lptr(1, 0);
  std::vector<int32_t> lcol, lrow;
  std::vector<uint8_t> lval;
  for (int64_t r : lrows) {
    for (int64_t k = 0; k < len[r]; ++k) {
      int64_t c; double v;
      src.entry(pos[r], k, c, v);
      lcol.push_back((int32_t)c);
      pack_values(lval, vt, v);
    }
    lptr.push_back(lptr.back() + len[r]);
    lrow.push_back((int32_t)r);
  }
  std::vector<uint8_t> val;
  val.reserve(dval.size() * value_bytes(vt));
  for (double v : dval) pack_values(val, vt, v);
  S.col.upload(col);
  S.val.upload(val);
  S.meta.upload(meta);
  S.tasks.upload(tasks);
  S.lptr.upload(lptr);
  S.lcol.upload(lcol);
  S.lval.upload(lval);
  S.lrow.upload(lrow);
  if (P > 1) S.part.alloc((size_t)P * n * sizeof(double));
  S.P = P;
  S.nlong = (int)lrows.size();
  S.ntasks = (int)tasks.size();
  M.slots = (int64_t)col.size() + (int64_t)lcol.size();
  M.dev_bytes = S.col.bytes + S.val.bytes + S.meta.bytes + S.tasks.bytes + S.lptr.bytes + S.lcol.bytes +
                S.lval.bytes + S.lrow.bytes + S.part.bytes;
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
  if (M.format == SPOPS_FORMAT_CSR && M.sweep.H == 0) {
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
  } else if (M.sweep.H > 0) {
    const auto &W = M.sweep;
    if (M.n_rows == 0) return;
    SweepArgs wa;
    wa.word = W.word.as<uint32_t>();
    wa.val = W.val.p;
    wa.gdesc = W.gdesc.as<int>();
    wa.wgrp = W.wgrp.as<int>();
    wa.wblk = W.wblk.as<int2>();
    wa.binfo = W.binfo.as<int4>();
    wa.rowidx = W.split > 0 ? W.rowidx.as<int>() : nullptr;
    wa.part = W.part.p;
    wa.x = x;
    wa.y = y;
    wa.n_rows = (int)M.n_rows;
    wa.n_cols = (int)M.n_cols;
    wa.rb = W.rb;
    wa.pstride = W.split > 0 ? 0 : (int)M.n_rows;
    wa.alpha = alpha;
    wa.beta = beta;
    hipLaunchKernelGGL((k_sweep_spmv<VT>), dim3((unsigned)W.nwg), dim3(kSweepBlock), 0, s, wa);
    if (W.split > 0 && W.nrtasks > 0) {
      SPOPS_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL((k_sweep_gather<typename T::X>), dim3((unsigned)W.nrtasks), dim3(kBlock), 0, s,
                         W.rtasks.as<int4>(), W.binfo.as<int4>(), W.part.as<typename T::X>(), W.rowidx.as<int>(), yy,
                         alpha, beta);
    } else if (W.split == 0 && W.H > 1) {
      SPOPS_HIP_CHECK(hipGetLastError());
      if (std::is_same<typename T::X, float>::value && M.n_rows % 4 == 0 && ((uintptr_t)W.part.p % 16) == 0 &&
          ((uintptr_t)y % 16) == 0) {
        const int n4 = (int)(M.n_rows / 4);
        hipLaunchKernelGGL(k_sweep_reduce4, dim3((unsigned)((n4 + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           W.part.as<f32x4>(), W.H, n4, static_cast<f32x4 *>(y), alpha, beta);
      } else {
        hipLaunchKernelGGL((k_sweep_reduce<typename T::X>), dim3((unsigned)((M.n_rows + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, W.part.as<typename T::X>(), W.H, (int)M.n_rows, yy, alpha, beta);
      }
    }
  } else if (M.format == SPOPS_FORMAT_SELL || M.format == SPOPS_FORMAT_ELL) {
    const auto &S = M.sell;
    SliceArgs sa;
    sa.col = S.col.as<int>();
    sa.val = S.val.p;
    sa.meta = S.meta.as<uint32_t>();
    sa.tasks = S.tasks.as<int4>();
    sa.lptr = S.lptr.as<long>();
    sa.lcol = S.lcol.as<int>();
    sa.lval = S.lval.p;
    sa.lrow = S.lrow.as<int>();
    sa.part = S.part.as<double>();
    sa.x = x;
    sa.y = y;
    int wg = 0;
    for (int q = 0; q < kXcds; ++q) {
      sa.sbase[q] = (long)S.sbase[q];
      sa.qoff[q] = S.qoff[q];
      wg = std::max(wg, (S.qoff[q + 1] - S.qoff[q] + kBlock / kWave - 1) / (kBlock / kWave));
    }
    sa.qoff[kXcds] = S.qoff[kXcds];
    sa.n_rows = (int)M.n_rows;
    sa.n_cols = (int)M.n_cols;
    sa.P = S.P;
    sa.nlong = S.nlong;
    sa.nlong_pad = (S.nlong + kXcds - 1) / kXcds * kXcds;
    sa.alpha = alpha;
    sa.beta = beta;
    const int grid = sa.nlong_pad + kXcds * wg;
    if (M.n_rows == 0) return;
    if (S.P == 1) {
      if (grid > 0) hipLaunchKernelGGL((k_slice_spmv<VT, true>), dim3(grid), dim3(kBlock), 0, s, sa);
    } else {
      if (grid > 0) hipLaunchKernelGGL((k_slice_spmv<VT, false>), dim3(grid), dim3(kBlock), 0, s, sa);
      SPOPS_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL((k_slice_reduce<typename T::X>), dim3((unsigned)((M.n_rows + kBlock - 1) / kBlock)),
                         dim3(kBlock), 0, s, S.part.as<double>(), S.P, (int)M.n_rows, yy, alpha, beta);
    }
  } else {
    const auto &B = M.bsr;
    int grid = (int)((B.mb + 3) / 4);
    if (grid == 0) return;
    auto mfma = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, B.brp.as<int>(), B.bcol.as<int>(),
                         B.data.as<typename T::V>(), (int)B.mb, (int)B.bR, (int)B.bC, (int)M.n_rows, (int)M.n_cols,
                         xx, yy, alpha, beta);
    };
    if (B.mfma && B.bR <= 16) mfma(k_bsr_mfma<VT, 1>);
    else if (B.mfma && B.bR <= 32) mfma(k_bsr_mfma<VT, 2>);
    else if (B.mfma) mfma(k_bsr_mfma<VT, kBsrMfmaMaxRows / 16>);
    else
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
  int64_t entry_index(int64_t g, int64_t k) const override { return rp[perm[g]] + k; }
};

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

// CSR handles: the swept layout beside the CSR arrays (SpMV only), chosen as for SELL.
static void add_csr_sweep(spops_matrix_s &M, const std::vector<int64_t> &rp, const std::vector<int64_t> &ci,
                          const void *values, int in_vt, int vt, int64_t mode) {
  const int64_t n = M.n_rows;
  if (mode == 1 || n == 0 || M.nnz == 0) return;
  std::vector<int64_t> ident(n);
  std::iota(ident.begin(), ident.end(), 0);
  CsrRows src(rp, ci, ident, values, in_vt);
  std::vector<int64_t> len(n);
  for (int64_t r = 0; r < n; ++r) len[r] = rp[r + 1] - rp[r];
  if (mode != 2 && !sweep_wanted(M, slice_count(M, src, ident, len, vt), vt)) return;
  const int64_t slots = M.slots;
  build_sweep_layout(M, src, ident, len, vt, true);
  M.slots = slots;
}

}  // namespace

static int csr_create_common(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                             const void *rowptr, const void *col, const void *values, int in_vt, int vt,
                             int64_t spmv_layout) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    SPOPS_REQUIRE(spmv_layout >= 0 && spmv_layout <= 2, SPOPS_ERROR_INVALID_VALUE, "spmv_layout must be 0, 1 or 2");
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
    upload_csr_plan(P, rp, n_rows);
    M->slots = nnz;
    add_csr_sweep(*M, rp, ci, values, in_vt, vt, spmv_layout);
    M->dev_bytes = P.rowptr.bytes + P.col.bytes + P.val.bytes + P.tasks.bytes + P.longs.bytes +
                   (M->sweep.H > 0 ? M->dev_bytes : 0);
    M->algo_bytes = spmv_algo_bytes(*M, index_type == SPOPS_INDEX_I32 ? 4 : 8);
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_csr_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz, int index_type,
                                const void *rowptr, const void *col, const void *values, int in_vt, int vt) {
  return csr_create_common(A, n_rows, n_cols, nnz, index_type, rowptr, col, values, in_vt, vt, g_sweep_mode);
}

extern "C" int spops_csr_create_layout(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                                       int index_type, const void *rowptr, const void *col, const void *values,
                                       int in_vt, int vt, int spmv_layout) {
  return csr_create_common(A, n_rows, n_cols, nnz, index_type, rowptr, col, values, in_vt, vt, spmv_layout);
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
  CsrRows src(rp, ci, perm, values, in_vt);
  build_slice_layout(*M, src, vt);
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
    build_slice_layout(*M, src, vt);
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
    const bool mfma = bsr_use_mfma(bR, bC);
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
    M->bsr.mb = mb; M->bsr.bR = bR; M->bsr.bC = bC; M->bsr.nnzb = nnzb; M->bsr.mfma = mfma;
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
    if (A->format == SPOPS_FORMAT_CSR) {
      info[8] = A->csr.nlong; info[9] = A->csr.ntasks;
      if (A->sweep.H > 0) info[13] = A->sweep.ngroups;     // SpMV runs on the swept layout
    }
    if (A->format == SPOPS_FORMAT_SELL || A->format == SPOPS_FORMAT_ELL) {
      info[8] = A->sell.nlong; info[9] = A->sell.ntasks;
      info[10] = A->sell.C; info[11] = A->sell.sigma; info[12] = A->sell.P; info[14] = A->ref_nchunks;
      if (A->sweep.H > 0) { info[8] = 0; info[9] = A->sweep.ngroups; info[12] = 0; }
      if (A->format == SPOPS_FORMAT_ELL) info[13] = A->ell_width;
    }
    if (A->format == SPOPS_FORMAT_BSR) {
      info[12] = A->bsr.bR; info[13] = A->bsr.bC; info[14] = A->bsr.nnzb; info[9] = A->bsr.mfma ? 1 : 0;
    }
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_set_tuning(int key, int64_t value, int64_t *old) {
  return guarded([&]() {
    if (key == SPOPS_TUNE_SLICE_BYTES) {
      SPOPS_REQUIRE(value >= 64, SPOPS_ERROR_INVALID_VALUE, "slice bytes must be >= 64");
      if (old) *old = g_slice_bytes;
      g_slice_bytes = value;
    } else if (key == SPOPS_TUNE_SPMV_LAYOUT) {
      SPOPS_REQUIRE(value >= 0 && value <= 2, SPOPS_ERROR_INVALID_VALUE, "SpMV layout must be 0, 1 or 2");
      if (old) *old = g_sweep_mode;
      g_sweep_mode = value;
    } else if (key == SPOPS_TUNE_SWEEP_WORKGROUPS) {
      SPOPS_REQUIRE(value >= 1, SPOPS_ERROR_INVALID_VALUE, "workgroup target must be >= 1");
      if (old) *old = g_sweep_workgroups;
      g_sweep_workgroups = value;
    } else if (key == SPOPS_TUNE_SWEEP_MAX_ROWS) {
      SPOPS_REQUIRE(value >= 1 && value <= kSweepMaxRows, SPOPS_ERROR_INVALID_VALUE,
                    "swept row block size must be in [1, %d]", kSweepMaxRows);
      if (old) *old = g_sweep_max_rows;
      g_sweep_max_rows = value;
    } else if (key == SPOPS_TUNE_SWEEP_SPLIT) {
      SPOPS_REQUIRE(value >= 0, SPOPS_ERROR_INVALID_VALUE, "swept split length must be >= 0");
      if (old) *old = g_sweep_split;
      g_sweep_split = value;
    } else if (key == SPOPS_TUNE_BSR_MFMA) {
      SPOPS_REQUIRE(value >= 0 && value <= 2, SPOPS_ERROR_INVALID_VALUE, "BSR MFMA mode must be 0, 1 or 2");
      if (old) *old = g_bsr_mfma;
      g_bsr_mfma = value;
    } else if (key == SPOPS_TUNE_TASK_PRODUCTS) {
      SPOPS_REQUIRE(value >= 0, SPOPS_ERROR_INVALID_VALUE, "task products must be >= 0");
      if (old) *old = g_task_products;
      g_task_products = value;
    } else if (key == SPOPS_TUNE_WIDE_CURSORS) {
      SPOPS_REQUIRE(value >= 0 && value <= 2, SPOPS_ERROR_INVALID_VALUE, "wide cursor mode must be 0, 1 or 2");
      if (old) *old = g_wide_cursors;
      g_wide_cursors = (int)value;
    } else if (key == SPOPS_TUNE_HEAVY_PRODUCTS) {
      SPOPS_REQUIRE(value >= 1, SPOPS_ERROR_INVALID_VALUE, "heavy product threshold must be >= 1");
      if (old) *old = g_heavy_products;
      g_heavy_products = value;
    } else if (key == SPOPS_TUNE_DENSE_PRODUCTS) {
      SPOPS_REQUIRE(value >= 1, SPOPS_ERROR_INVALID_VALUE, "dense product threshold must be >= 1");
      if (old) *old = g_dense_products;
      g_dense_products = value;
    } else {
      throw_error(SPOPS_ERROR_INVALID_VALUE, "unknown tuning key %d", key);
    }
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_matrix_layout(spops_matrix A, int which, void *host_out) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr && host_out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null argument");
    SPOPS_REQUIRE(A->format == SPOPS_FORMAT_SELL ||
                      (which >= 2 && (A->format == SPOPS_FORMAT_ELL || A->format == SPOPS_FORMAT_CSR)),
                  SPOPS_ERROR_NOT_SUPPORTED, "layout arrays exist for SELL only");
    if (which == 0) std::memcpy(host_out, A->sell.ref_reorder.data(), A->sell.ref_reorder.size() * 4);
    else if (which == 1) std::memcpy(host_out, A->sell.ref_chunk_len.data(), A->sell.ref_chunk_len.size() * 4);
    else if (which == 2) {
      const int64_t d[4] = {A->sweep.rb, A->sweep.H, A->sweep.n_rb, A->sweep.ngroups};
      std::memcpy(host_out, d, sizeof(d));
    } else if (which == 3) {
      const int64_t d[4] = {A->sweep.split, A->sweep.nwg, A->sweep.plane_elems, A->sweep.nrtasks};
      std::memcpy(host_out, d, sizeof(d));
    }
    else throw_error(SPOPS_ERROR_INVALID_VALUE, "unknown layout array %d", which);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_matrix_destroy(spops_matrix A) {
  delete A;
  return SPOPS_SUCCESS;
}

extern "C" int spops_bsr_create_from_blocks(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t bR, int64_t bC,
                                            int64_t mb, int64_t nnzb, int index_type, const void *row_ptr,
                                            const void *block_colindex, const void *block_data, int in_vt, int vt) {
  return guarded([&]() {
    SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    check_vt(in_vt, vt);
    SPOPS_REQUIRE(bR > 0 && bC > 0 && mb == (n_rows + bR - 1) / bR, SPOPS_ERROR_INVALID_VALUE,
                  "mb must be ceil(n_rows / bR)");
    const bool mfma = bsr_use_mfma(bR, bC);
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
    M->bsr.mb = mb; M->bsr.bR = bR; M->bsr.bC = bC; M->bsr.nnzb = nnzb; M->bsr.mfma = mfma;
    M->slots = nnzb * bR * bC;
    M->dev_bytes = M->bsr.brp.bytes + M->bsr.bcol.bytes + M->bsr.data.bytes;
    int64_t sV = (int64_t)value_bytes(vt), sX = (int64_t)xy_bytes(vt), sI = index_type == SPOPS_INDEX_I32 ? 4 : 8;
    M->algo_bytes = (mb + 1) * sI + nnzb * sI + nnzb * bR * bC * sV + n_cols * sX + n_rows * sX;
    *A = M.release();
    return SPOPS_SUCCESS;
  });
}
