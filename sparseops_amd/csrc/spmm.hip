// SpMM, SDDMM, transpose and value updates on CSR matrices: the multi-RHS and the
// differentiable forms of SpOps_csr (src/spmv_csr.cpp:165-185 applied column by column).
//
//   Y = alpha * A * X + beta * Y        X: n_cols x k, Y: n_rows x k, both row-major
//   out[e] = sum_c G[row_e, c] * X[col_e, c]      (SDDMM on A's pattern: d(loss)/d(values))
//   A^T with the entry permutation kept, so that A^T's values follow A's (dX = A^T * G)
//
// X, Y, G in FP64, FP32 or BF16; accumulation in FP64 when the matrix or X is FP64, else
// FP32.  Kernels:
//   k_spmm_rows    : one row per group of L lanes (64 / L rows per wave), each lane covering
//                    VEC consecutive columns with one 16-byte load per entry when k and the
//                    alignment of X and Y allow it (else one column per lane), eight row
//                    entries in flight (four with 2-byte X); rows ordered by length inside windows of
//                    kSpmmWindow rows.  Rows longer than kSpmmLongRow entries are left to
//   k_spmm_seg     : one workgroup per segment of kSpmmSegment entries of a long row, its
//                    partial row of sums to a scratch buffer, then
//   k_spmm_seg_sum : one workgroup per long row adds its segments in order (deterministic).
//   k_sddmm_rows   : a group of L lanes per chunk of at most 64 entries of one row: the row
//                    of G loaded once, then one 16-byte load of X's row per entry and lane,
//                    cross-lane sum (k a multiple of the vector, at most 64 vectors);
//   k_sddmm        : otherwise entry-parallel, a group of L lanes per entry with loads of
//                    both the G and the X row, cross-lane sum.
//   k_spmm_sddmm_rows / k_spmm_sddmm_seg : the differentiable backward in one pass over a
//                    transposed handle: k_spmm_rows / k_spmm_seg for dX = A^T G plus each entry's
//                    dot product with the row of X (loaded once per row of A^T), written in
//                    A^T's order and gathered into A's by the inverse permutation.
#include <algorithm>
#include <memory>
#include <mutex>
#include <numeric>

#include "common.hpp"
#include "spmv_internal.hpp"

namespace spops {

namespace {

template <int C> struct Ty;
template <> struct Ty<SPOPS_VALUE_F64> {
  using T = double;
  static __device__ __forceinline__ double ld(double v) { return v; }
  static __device__ __forceinline__ double st(double d) { return d; }
};
template <> struct Ty<SPOPS_VALUE_F32> {
  using T = float;
  static __device__ __forceinline__ double ld(float v) { return (double)v; }
  static __device__ __forceinline__ float st(double d) { return (float)d; }
};
template <> struct Ty<SPOPS_VALUE_BF16> {
  using T = uint16_t;
  static __device__ __forceinline__ double ld(uint16_t v) { return (double)bf16_to_float(v); }
  // double -> float with round-to-odd, then float -> bf16 round to nearest even: a direct
  // double -> bf16 rounding (the device twin of double_to_bf16_rne)
  static __device__ __forceinline__ uint16_t st(double d) {
    float f = (float)d;
    if ((double)f != d && f != 0.0f && f == f && f - f == 0.0f) {
      uint32_t o = __float_as_uint(f);
      if ((o & 1u) == 0u) o = (((double)f < d) == (f > 0.0f)) ? o + 1u : o - 1u;
      f = __uint_as_float(o);
    }
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
};

template <int VC, int XC>
using AccT = typename std::conditional<VC == SPOPS_VALUE_F64 || XC == SPOPS_VALUE_F64, double, float>::type;

// VEC consecutive elements of one row of X, Y or G as one aligned vector access
template <int C, int VEC> struct Vec {
  using T = typename Ty<C>::T;
  struct alignas(VEC * sizeof(T)) V { T e[VEC]; };
};

// Rows of at most kSpmmLongRow entries, taken in `order` (by length inside windows of rows,
// so the rows of a wave have about the same length).  A row is served by a group of L lanes
// (64 / L rows per wave); lane q of the group covers columns c .. c + VEC - 1 with
// c = cb + q VEC, loading them as one vector per entry.  U entries of the row are in flight.
template <int VC, int XC, int YC, typename RP, int VEC, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_rows(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                      const typename Ty<VC>::T *__restrict__ val,
                                                      const int *__restrict__ vperm,
                                                      const int *__restrict__ order, int64_t n_short, int k, int L,
                                                      const typename Ty<XC>::T *__restrict__ X,
                                                      typename Ty<YC>::T *__restrict__ Y, double alpha, double beta) {
  using Acc = AccT<VC, XC>;
  using XV = typename Vec<XC, VEC>::V;
  using YV = typename Vec<YC, VEC>::V;
  const int lane = lane_id(), R = kWave / L, g = lane / L, q = lane % L;
  const int64_t slot = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * R + g;
  if (slot >= n_short) return;                      // lane groups past the last row
  const int64_t row = order[slot];
  const int64_t p0 = (int64_t)rowptr[row], p1 = (int64_t)rowptr[row + 1];
  for (int cb = 0; cb < k; cb += L * VEC) {
    const int c = cb + q * VEC;
    const bool cv = c < k;                          // k % VEC == 0: a vector is all in or all out
    Acc acc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[e] = 0;
    for (int64_t p = p0; p < p1; p += U) {
      int cc[U];
      Acc vv[U];
      XV xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = p + u < p1;
        cc[u] = ok ? col[p + u] : 0;
        vv[u] = ok ? (Acc)Ty<VC>::ld(val[vperm ? (int64_t)vperm[p + u] : p + u]) : Acc(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        XV t{};
        if (cv && p + u < p1) t = *reinterpret_cast<const XV *>(X + (int64_t)cc[u] * k + c);
        xv[u] = t;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = vv[u] * (Acc)Ty<XC>::ld(xv[u].e[e]) + acc[e];
    }
    if (cv) {
      YV *dst = reinterpret_cast<YV *>(Y + row * k + c);
      YV o;
      if (beta != 0.0) {
        const YV old = *dst;
#pragma unroll
        for (int e = 0; e < VEC; ++e) o.e[e] = Ty<YC>::st(alpha * (double)acc[e] + beta * Ty<YC>::ld(old.e[e]));
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) o.e[e] = Ty<YC>::st(alpha * (double)acc[e]);
      }
      *dst = o;
    }
  }
}

// Long rows, pass 1: one workgroup per segment of kSpmmSegment entries of a long row.  The
// wave's 64 / L entry slots and the 4 waves each take every (4 * 64 / L)-th entry; the slots
// are summed with cross-lane adds, the waves through LDS, and the segment's row of partial
// sums is written to part[segment * k ...].
template <int VC, int XC, typename RP, int VEC, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_seg(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                     const typename Ty<VC>::T *__restrict__ val,
                                                     const int *__restrict__ vperm,
                                                     const int *__restrict__ long_rows, const int2 *__restrict__ segs,
                                                     int k, int L, const typename Ty<XC>::T *__restrict__ X,
                                                     AccT<VC, XC> *__restrict__ part) {
  using Acc = AccT<VC, XC>;
  using XV = typename Vec<XC, VEC>::V;
  constexpr int W = kBlock / kWave;
  __shared__ Acc red[W][kWave * VEC];
  const int2 sg = segs[blockIdx.x];
  const int64_t row = long_rows[sg.x];
  const int64_t q0 = (int64_t)rowptr[row] + sg.y;
  const int64_t q1 = std::min<int64_t>((int64_t)rowptr[row + 1], q0 + kSpmmSegment);
  const int lane = lane_id(), w = wave_id(), E = kWave / L, es = lane / L, q = lane % L;
  const int64_t stride = (int64_t)W * E;
  for (int cb = 0; cb < k; cb += L * VEC) {
    const int c = cb + q * VEC;
    const bool cv = c < k;
    Acc acc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[e] = 0;
    for (int64_t p = q0 + (int64_t)w * E + es; p < q1; p += stride * U) {
      int cc[U];
      Acc vv[U];
      XV xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = p + u * stride < q1;
        const int64_t e = p + u * stride;
        cc[u] = ok ? col[e] : 0;
        vv[u] = ok ? (Acc)Ty<VC>::ld(val[vperm ? (int64_t)vperm[e] : e]) : Acc(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        XV t{};
        if (cv && p + u * stride < q1) t = *reinterpret_cast<const XV *>(X + (int64_t)cc[u] * k + c);
        xv[u] = t;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = vv[u] * (Acc)Ty<XC>::ld(xv[u].e[e]) + acc[e];
    }
    for (int m = L; m < kWave; m <<= 1)
#pragma unroll
      for (int e = 0; e < VEC; ++e) acc[e] += __shfl_xor(acc[e], m, kWave);
    if (es == 0)
#pragma unroll
      for (int e = 0; e < VEC; ++e) red[w][q * VEC + e] = acc[e];
    __syncthreads();
    if (w == 0 && es == 0 && cv) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        Acc t = red[0][q * VEC + e];
#pragma unroll
        for (int v = 1; v < W; ++v) t += red[v][q * VEC + e];
        part[(int64_t)blockIdx.x * k + c + e] = t;
      }
    }
    __syncthreads();
  }
}

// Long rows, pass 2: one workgroup per long row sums its segments in order (deterministic)
// and writes y = alpha * sum + beta * y.
template <typename Acc, int YC>
__global__ __launch_bounds__(kBlock) void k_spmm_seg_sum(const int *__restrict__ long_rows,
                                                         const int *__restrict__ segoff, const Acc *__restrict__ part,
                                                         int k, typename Ty<YC>::T *__restrict__ Y, double alpha,
                                                         double beta) {
  const int li = blockIdx.x;
  const int64_t row = long_rows[li];
  const int s0 = segoff[li], s1 = segoff[li + 1];
  for (int c = threadIdx.x; c < k; c += kBlock) {
    Acc t = 0;
    for (int sg = s0; sg < s1; ++sg) t += part[(int64_t)sg * k + c];
    double o = alpha * (double)t;
    if (beta != 0.0) o += beta * Ty<YC>::ld(Y[row * k + c]);
    Y[row * k + c] = Ty<YC>::st(o);
  }
}

// Sums U per-lane partials (one per entry of a step) over the L lanes of each lane group (L a
// power of two, groups aligned to L).  Each step pairs lanes q and q ^ m and halves the entries
// a lane holds: the lane keeps the lower or upper half (by bit m of q) and adds the partner's
// values of that half, so the U sums take U - 1 + log2(L / U) shuffles (instead of U log2 L).
// Afterwards lane q holds the complete sums of entries base + i in part[i], i < U / min(L, U);
// the return value is base.  Lanes with q >= U hold copies.
template <int U, typename A>
__device__ __forceinline__ int group_sum_entries(A (&part)[U], int q, int L) {
  int base = 0;
#pragma unroll
  for (int s = U / 2, m = 1; s >= 1; s >>= 1, m <<= 1) {
    if (m < L) {
      const bool hi = (q & m) != 0;
#pragma unroll
      for (int i = 0; i < s; ++i) {
        const A send = hi ? part[i] : part[i + s];
        const A keep = hi ? part[i + s] : part[i];
        part[i] = keep + __shfl_xor(send, m, kWave);
      }
      if (hi) base += s;
    }
  }
  for (int m = U; m < L; m <<= 1) part[0] += __shfl_xor(part[0], m, kWave);
  return base;
}

// out[p + base + i] = part[i] for the entries a lane holds after group_sum_entries (p + u < p1)
template <int U, int OC, typename A>
__device__ __forceinline__ void group_write_entries(const A (&part)[U], int base, int q, int L, int64_t p, int64_t p1,
                                                    int64_t stride, typename Ty<OC>::T *__restrict__ out) {
  if (q >= U) return;
  const int held = L >= U ? 1 : U / L;
#pragma unroll
  for (int i = 0; i < U; ++i) {
    const int64_t e = p + (int64_t)(base + i) * stride;
    if (i < held && e < p1) out[e] = Ty<OC>::st((double)part[i]);
  }
}

// <a, b> over VEC BF16 elements with v_dot2c_f32_bf16 (two products and an FP32 addend per
// instruction, each product exact in FP32), in two chains; the operands stay packed, so no
// conversion of either row to FP32 is needed for the dot product.
typedef __bf16 __attribute__((ext_vector_type(2))) bf16x2_t;
__device__ __forceinline__ bf16x2_t bf16_pair(uint16_t lo, uint16_t hi) {
  return __builtin_bit_cast(bf16x2_t, (uint32_t)lo | ((uint32_t)hi << 16));
}
template <int VEC>
__device__ __forceinline__ float dot_bf16(const uint16_t (&a)[VEC], const uint16_t (&b)[VEC]) {
  static_assert(VEC % 4 == 0, "pairs of pairs");
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int i = 0; i < VEC; i += 4) {
    s0 = __builtin_amdgcn_fdot2_f32_bf16(bf16_pair(a[i], a[i + 1]), bf16_pair(b[i], b[i + 1]), s0, false);
    s1 = __builtin_amdgcn_fdot2_f32_bf16(bf16_pair(a[i + 2], a[i + 3]), bf16_pair(b[i + 2], b[i + 3]), s1, false);
  }
  return s0 + s1;
}

// The backward of the differentiable SpMM in one pass over the transposed handle AT (its rows
// are the source matrix's columns): dX = AT G as k_spmm_rows computes it (alpha 1, beta 0), and
// for each entry e of AT, dT[e] = <G[col_e, :], X[row, :]>, the SDDMM of the source matrix's
// entry tperm[e], in AT's entry order (a gather through the inverse permutation puts it in the
// source's order).  Each gathered row of G serves both products, and the row of X is loaded
// once per AT row, so the SDDMM's gathers of X rows disappear.  One lane group covers all k
// columns (L VEC >= k); the U dot products of a step are summed by group_sum_entries.
template <int VC, int GC, int XC, typename RP, int VEC, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_sddmm_rows(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                            const typename Ty<VC>::T *__restrict__ val,
                                                            const int *__restrict__ order, int64_t n_short, int k, int L,
                                                            const typename Ty<GC>::T *__restrict__ G,
                                                            const typename Ty<XC>::T *__restrict__ X,
                                                            typename Ty<XC>::T *__restrict__ dX,
                                                            typename Ty<VC>::T *__restrict__ dT) {
  using Acc = AccT<VC, GC>;
  using Ad = AccT<GC, XC>;
  using GV = typename Vec<GC, VEC>::V;
  using XV = typename Vec<XC, VEC>::V;
  const int lane = lane_id(), R = kWave / L, g = lane / L, q = lane % L;
  const int64_t slot = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * R + g;
  if (slot >= n_short) return;                      // lane groups past the last row
  const int64_t row = order[slot];
  const int64_t p0 = (int64_t)rowptr[row], p1 = (int64_t)rowptr[row + 1];
  const int c = q * VEC;
  const bool cv = c < k;
  // BF16 G and X: dot products on the packed pairs (v_dot2c_f32_bf16)
  constexpr bool kDot2 = GC == SPOPS_VALUE_BF16 && XC == SPOPS_VALUE_BF16 && VEC % 4 == 0;
  Ad xr[VEC];
  XV xraw{};
  if (cv) xraw = *reinterpret_cast<const XV *>(X + row * k + c);
#pragma unroll
  for (int e = 0; e < VEC; ++e) xr[e] = (Ad)Ty<XC>::ld(xraw.e[e]);
  Acc acc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) acc[e] = 0;
  for (int64_t p = p0; p < p1; p += U) {
    int cc[U];
    Acc vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = p + u < p1;
      cc[u] = ok ? col[p + u] : 0;
      vv[u] = ok ? (Acc)Ty<VC>::ld(val[p + u]) : Acc(0);
    }
    GV gv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      GV t{};
      if (cv && p + u < p1) t = *reinterpret_cast<const GV *>(G + (int64_t)cc[u] * k + c);
      gv[u] = t;
    }
    Ad part[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (kDot2) {
        part[u] = dot_bf16<VEC>(gv[u].e, xraw.e);
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = vv[u] * (Acc)Ty<GC>::ld(gv[u].e[e]) + acc[e];
      } else {
        part[u] = 0;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          acc[e] = vv[u] * (Acc)Ty<GC>::ld(gv[u].e[e]) + acc[e];
          part[u] += (Ad)Ty<GC>::ld(gv[u].e[e]) * xr[e];
        }
      }
    }
    const int base = group_sum_entries<U>(part, q, L);
    group_write_entries<U, VC>(part, base, q, L, p, p1, 1, dT);
  }
  if (cv) {
    XV o;
#pragma unroll
    for (int e = 0; e < VEC; ++e) o.e[e] = Ty<XC>::st((double)acc[e]);
    *reinterpret_cast<XV *>(dX + row * k + c) = o;
  }
}

// The same for the segments of AT's long rows (k_spmm_seg's partition of the entries): the
// segment's partial row of dX to part[], and each entry's dot product with the row of X
// (loaded once per segment) summed over its slot's L lanes and written to dT[e], U entries per
// slot and step.
template <int VC, int GC, int XC, typename RP, int VEC, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_sddmm_seg(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                           const typename Ty<VC>::T *__restrict__ val,
                                                           const int *__restrict__ long_rows,
                                                           const int2 *__restrict__ segs, int k, int L,
                                                           const typename Ty<GC>::T *__restrict__ G,
                                                           const typename Ty<XC>::T *__restrict__ X,
                                                           AccT<VC, GC> *__restrict__ part,
                                                           typename Ty<VC>::T *__restrict__ dT) {
  using Acc = AccT<VC, GC>;
  using Ad = AccT<GC, XC>;
  using GV = typename Vec<GC, VEC>::V;
  using XV = typename Vec<XC, VEC>::V;
  constexpr int W = kBlock / kWave;
  __shared__ Acc red[W][kWave * VEC];
  const int2 sg = segs[blockIdx.x];
  const int64_t row = long_rows[sg.x];
  const int64_t q0 = (int64_t)rowptr[row] + sg.y;
  const int64_t q1 = std::min<int64_t>((int64_t)rowptr[row + 1], q0 + kSpmmSegment);
  const int lane = lane_id(), w = wave_id(), E = kWave / L, es = lane / L, q = lane % L;
  const int64_t stride = (int64_t)W * E;
  const int c = q * VEC;
  const bool cv = c < k;
  // BF16 G and X: dot products on the packed pairs (v_dot2c_f32_bf16)
  constexpr bool kDot2 = GC == SPOPS_VALUE_BF16 && XC == SPOPS_VALUE_BF16 && VEC % 4 == 0;
  Ad xr[VEC];
  XV xraw{};
  if (cv) xraw = *reinterpret_cast<const XV *>(X + row * k + c);
#pragma unroll
  for (int e = 0; e < VEC; ++e) xr[e] = (Ad)Ty<XC>::ld(xraw.e[e]);
  Acc acc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) acc[e] = 0;
  for (int64_t p = q0 + (int64_t)w * E + es; p < q1; p += stride * U) {
    int cc[U];
    Acc vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = p + u * stride;
      const bool ok = e < q1;
      cc[u] = ok ? col[e] : 0;
      vv[u] = ok ? (Acc)Ty<VC>::ld(val[e]) : Acc(0);
    }
    GV gv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      GV t{};
      if (cv && p + u * stride < q1) t = *reinterpret_cast<const GV *>(G + (int64_t)cc[u] * k + c);
      gv[u] = t;
    }
    Ad pd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (kDot2) {
        pd[u] = dot_bf16<VEC>(gv[u].e, xraw.e);
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = vv[u] * (Acc)Ty<GC>::ld(gv[u].e[e]) + acc[e];
      } else {
        pd[u] = 0;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          acc[e] = vv[u] * (Acc)Ty<GC>::ld(gv[u].e[e]) + acc[e];
          pd[u] += (Ad)Ty<GC>::ld(gv[u].e[e]) * xr[e];
        }
      }
    }
    const int base = group_sum_entries<U>(pd, q, L);
    group_write_entries<U, VC>(pd, base, q, L, p, q1, stride, dT);
  }
  for (int m = L; m < kWave; m <<= 1)
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[e] += __shfl_xor(acc[e], m, kWave);
  if (es == 0)
#pragma unroll
    for (int e = 0; e < VEC; ++e) red[w][q * VEC + e] = acc[e];
  __syncthreads();
  if (w == 0 && es == 0 && cv) {
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      Acc t = red[0][q * VEC + e];
#pragma unroll
      for (int v = 1; v < W; ++v) t += red[v][q * VEC + e];
      part[(int64_t)blockIdx.x * k + c + e] = t;
    }
  }
}

// The two passes of a PermPlan (bit copies of T-sized elements, tiles of TS elements: TS * sizeof(T)
// bytes of LDS).  A thread handles elements t + i kPermThreads of the tile (every load
// instruction coalesced), sixteen at a time: all sixteen loads are issued before their LDS
// stores, so a thread keeps sixteen loads in flight instead of one (at R-MAT 20 pass 1 took 91 us
// with batches of eight against 74 us, pass 2 303 us with one load per thread in each step's
// chain against 116 us).  The tile's LDS leaves one workgroup per CU, so it has kPermThreads =
// 1024 threads (16 waves) to keep enough loads in flight: 256 threads took 74.8 / 117.0 us for
// the two passes, 512 61.9 / 97.1, 1024 58.3 / 93.6.  Pass 1: workgroup t takes
// sources t TS.., orders them in LDS by destination tile and writes the ordered tile back to
// the same range of the staging array (coalesced).
constexpr int kPermBatch = 16;
constexpr int kPermThreads = 1024;

template <typename T, int TS>
__global__ __launch_bounds__(kPermThreads) void k_perm_tiles(const T *__restrict__ x, const uint16_t *__restrict__ loc1,
                                                       int64_t n, T *__restrict__ stg) {
  constexpr int B = kPermBatch;
  static_assert(TS % (kPermThreads * B) == 0, "whole batches per tile");
  __shared__ T lds[TS];
  const int64_t j0 = (int64_t)blockIdx.x * TS;
  const int cnt = (int)std::min<int64_t>(TS, n - j0);
  // 32-bit offsets into buffer resources over the tile (offsets past it load 0, not stored)
  const rsrc_t rx = make_rsrc(x + j0, (uint64_t)cnt * sizeof(T));
  const rsrc_t rl = make_rsrc(loc1 + j0, (uint64_t)cnt * 2);
  const int t = (int)threadIdx.x;
  for (int q0 = 0; q0 < TS; q0 += kPermThreads * B) {
    T v[B];
    uint16_t l[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t q = (uint32_t)(q0 + t + i * kPermThreads);
      v[i] = bload<T>(rx, q * (uint32_t)sizeof(T));
      l[i] = bload<uint16_t>(rl, q * 2u);
    }
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (q0 + t + i * kPermThreads < cnt) lds[l[i]] = v[i];
  }
  __syncthreads();
  T *ss = stg + (int64_t)blockIdx.x * (TS + kPermPad);
  for (int q = t; q < cnt; q += kPermThreads) ss[q] = lds[q];
}

// Pass 2: workgroup b takes destination tile b: its elements, which lie in every staged source
// tile as one run (src2: their staging addresses, run after run), to their places in LDS
// (loc2), then the tile out (coalesced).
template <typename T, int TS>
__global__ __launch_bounds__(kPermThreads) void k_perm_buckets(const T *__restrict__ stg, int64_t n_stage,
                                                         const int *__restrict__ src2,
                                                         const uint16_t *__restrict__ loc2, int64_t n,
                                                         T *__restrict__ y) {
  // per batch: the addresses, then the data loads, then the LDS stores (32-bit offsets into
  // buffer resources: the staging array stays below 2 GiB; offsets past the tile load 0 and
  // are not stored)
  constexpr int B = kPermBatch;
  static_assert(TS % (kPermThreads * B) == 0, "whole batches per tile");
  __shared__ T lds[TS];
  const int64_t k0 = (int64_t)blockIdx.x * TS;
  const int cnt = (int)std::min<int64_t>(TS, n - k0);
  const rsrc_t ra = make_rsrc(src2 + k0, (uint64_t)cnt * 4);
  const rsrc_t rl = make_rsrc(loc2 + k0, (uint64_t)cnt * 2);
  const rsrc_t rs = make_rsrc(stg, (uint64_t)n_stage * sizeof(T));
  const int t = (int)threadIdx.x;
  for (int q0 = 0; q0 < TS; q0 += kPermThreads * B) {
    uint32_t a[B];
    uint16_t l[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t q = (uint32_t)(q0 + t + i * kPermThreads);
      a[i] = bload<uint32_t>(ra, q * 4u);
      l[i] = bload<uint16_t>(rl, q * 2u);
    }
    T v[B];
#pragma unroll
    for (int i = 0; i < B; ++i) v[i] = bload<T>(rs, a[i] * (uint32_t)sizeof(T));
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (q0 + t + i * kPermThreads < cnt) lds[l[i]] = v[i];
  }
  __syncthreads();
  T *ys = y + k0;
  for (int q = t; q < cnt; q += kPermThreads) ys[q] = lds[q];
}

// out[e] = sum_c G[row_e, c] * X[col_e, c], entry-parallel: a group of L lanes per entry
// (64 / L entries per wave step, U steps per wave), lane q of the group covering columns
// c .. c + VEC - 1 (c = cb + q VEC) with one vector load from G's row and one from X's row;
// the group is summed with cross-lane adds.  erow holds each entry's row.
template <int GC, int XC, int OC, int VEC>
__global__ __launch_bounds__(kBlock) void k_sddmm(const int *__restrict__ erow, const int *__restrict__ col,
                                                  int64_t nnz, int k, int L, const typename Ty<GC>::T *__restrict__ G,
                                                  const typename Ty<XC>::T *__restrict__ X,
                                                  typename Ty<OC>::T *__restrict__ out) {
  using Acc = AccT<GC, XC>;
  using GV = typename Vec<GC, VEC>::V;
  using XV = typename Vec<XC, VEC>::V;
  constexpr int U = 4;
  const int lane = lane_id(), E = kWave / L, es = lane / L, q = lane % L;
  const int64_t base = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * E * U + es;
  int rr[U], cc[U];
  bool ok[U];
  Acc part[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t p = base + (int64_t)u * E;
    ok[u] = p < nnz;
    rr[u] = ok[u] ? erow[p] : 0;
    cc[u] = ok[u] ? col[p] : 0;
    part[u] = 0;
  }
  for (int cb = 0; cb < k; cb += L * VEC) {
    const int c = cb + q * VEC;
    const bool cv = c < k;
    GV gv[U];
    XV xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      GV a{};
      XV b{};
      if (cv && ok[u]) {
        a = *reinterpret_cast<const GV *>(G + (int64_t)rr[u] * k + c);
        b = *reinterpret_cast<const XV *>(X + (int64_t)cc[u] * k + c);
      }
      gv[u] = a;
      xv[u] = b;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < VEC; ++e) part[u] += (Acc)Ty<GC>::ld(gv[u].e[e]) * (Acc)Ty<XC>::ld(xv[u].e[e]);
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    for (int m = 1; m < L; m <<= 1) part[u] += __shfl_xor(part[u], m, kWave);
  if (q == 0) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) out[base + (int64_t)u * E] = Ty<OC>::st((double)part[u]);
  }
}

// The same by rows: a group of L lanes per task, a chunk of at most kSddmmChunk entries of one
// row (tasks {row, first entry in the row}, the rows of a window longest first).  The group
// loads its row of G once (lane q: columns q VEC .. q VEC + VEC - 1, all of k when
// k <= L VEC), then per entry only the row of X; U entries in flight, summed across the group
// by group_sum_entries (each lane q < U then writes the entries it holds).
template <int GC, int XC, int OC, typename RP, int VEC, int U>
__global__ __launch_bounds__(kBlock) void k_sddmm_rows(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                       const int2 *__restrict__ tasks, int64_t ntasks, int k, int L,
                                                       const typename Ty<GC>::T *__restrict__ G,
                                                       const typename Ty<XC>::T *__restrict__ X,
                                                       typename Ty<OC>::T *__restrict__ out) {
  using Acc = AccT<GC, XC>;
  using GV = typename Vec<GC, VEC>::V;
  using XV = typename Vec<XC, VEC>::V;
  const int lane = lane_id(), R = kWave / L, g = lane / L, q = lane % L;
  const int64_t slot = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * R + g;
  if (slot >= ntasks) return;                       // lane groups past the last task
  const int2 t = tasks[slot];
  const int64_t row = t.x;
  const int64_t p0 = (int64_t)rowptr[row] + t.y;
  const int64_t p1 = std::min<int64_t>((int64_t)rowptr[row + 1], p0 + kSddmmChunk);
  const int c = q * VEC;
  const bool cv = c < k;
  // BF16 G and X: dot products on the packed pairs (v_dot2c_f32_bf16)
  constexpr bool kDot2 = GC == SPOPS_VALUE_BF16 && XC == SPOPS_VALUE_BF16 && VEC % 4 == 0;
  Acc gr[VEC];
  GV graw{};
  if (cv) graw = *reinterpret_cast<const GV *>(G + row * k + c);
#pragma unroll
  for (int e = 0; e < VEC; ++e) gr[e] = (Acc)Ty<GC>::ld(graw.e[e]);
  for (int64_t p = p0; p < p1; p += U) {
    int cc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) cc[u] = p + u < p1 ? col[p + u] : 0;
    XV xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      XV t{};
      if (cv && p + u < p1) t = *reinterpret_cast<const XV *>(X + (int64_t)cc[u] * k + c);
      xv[u] = t;
    }
    Acc part[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (kDot2) {
        part[u] = dot_bf16<VEC>(graw.e, xv[u].e);
      } else {
        part[u] = 0;
#pragma unroll
        for (int e = 0; e < VEC; ++e) part[u] += gr[e] * (Acc)Ty<XC>::ld(xv[u].e[e]);
      }
    }
    const int base = group_sum_entries<U>(part, q, L);
    group_write_entries<U, OC>(part, base, q, L, p, p1, 1, out);
  }
}

// erow[e] = row of entry e; one wave per row
template <typename RP>
__global__ __launch_bounds__(kBlock) void k_entry_rows(const RP *__restrict__ rowptr, int64_t n_rows,
                                                       int *__restrict__ erow) {
  const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (row >= n_rows) return;
  for (int64_t p = (int64_t)rowptr[row] + lane_id(); p < (int64_t)rowptr[row + 1]; p += kWave) erow[p] = (int)row;
}

template <int SC, int DC, typename IDX>
__global__ __launch_bounds__(kBlock) void k_convert_values(const typename Ty<SC>::T *__restrict__ src,
                                                           const IDX *__restrict__ perm, int64_t n,
                                                           typename Ty<DC>::T *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dst[i] = Ty<DC>::st(Ty<SC>::ld(src[perm ? (int64_t)perm[i] : i]));
}

inline unsigned nblk(int64_t n, int per) { return grid_blocks(n, per); }

// swept-layout values from the CSR values (bit copies; padding slots stay zero)
template <typename B>
__global__ __launch_bounds__(kBlock) void k_sweep_refresh(const int32_t *__restrict__ src, int64_t slots,
                                                          const B *__restrict__ csr_val, B *__restrict__ sweep_val) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= slots) return;
  const int32_t e = src[i];
  sweep_val[i] = e >= 0 ? csr_val[e] : B(0);
}

void require_csr(const spops_matrix_s *A) {
  SPOPS_REQUIRE(A != nullptr, SPOPS_ERROR_INVALID_VALUE, "null matrix");
  SPOPS_REQUIRE(A->format == SPOPS_FORMAT_CSR, SPOPS_ERROR_NOT_SUPPORTED,
                "this operation is implemented for CSR matrices (create the matrix with spops_csr_create)");
}

void require_type(int t, const char *what) {
  SPOPS_REQUIRE(t == SPOPS_VALUE_F64 || t == SPOPS_VALUE_F32 || t == SPOPS_VALUE_BF16, SPOPS_ERROR_INVALID_VALUE,
                "%s type must be F64, F32 or BF16", what);
}

// ---- dispatch on (value, x, y) types and rowptr width -----------------------------------
inline bool aligned_to(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

// lanes per row (or entry) for k columns, VEC columns per lane: a power of two, at most 64
inline int lanes_for(int k, int vec) {
  const int need = (k + vec - 1) / vec;
  int L = 1;
  while (L < need && L < kWave) L <<= 1;
  return L;
}

template <int VC, int XC, int YC, typename RP, int VEC>
void spmm_launch_vec(const spops_matrix_s &A, int k, double alpha, const void *X, double beta, void *Y,
                     hipStream_t s, void *ext_part, const void *ext_val) {
  using Acc = AccT<VC, XC>;
  const auto &C = A.csr;
  const RP *rp = C.rowptr.as<RP>();
  const auto *v = ext_val ? static_cast<const typename Ty<VC>::T *>(ext_val) : C.val.as<typename Ty<VC>::T>();
  // caller's values of a transposed handle are in the source matrix's entry order
  const int *vp = ext_val && C.tperm.p ? C.tperm.as<int>() : nullptr;
  const auto *x = static_cast<const typename Ty<XC>::T *>(X);
  auto *y = static_cast<typename Ty<YC>::T *>(Y);
  const int L = lanes_for(k, VEC);
  // entries in flight per row: with vectors of eight 2-byte elements (eight FP32 sums per lane)
  // eight entries take 100 VGPRs (4 waves per SIMD) and four 64 (8 waves): R-MAT 20, k = 64,
  // BF16 X 0.81 -> 0.67 ms; FP32 / FP64 X keep eight (4 there were 1-5 % slower)
  constexpr int kRowSteps = VEC >= 8 ? 4 : 8, kSegSteps = VEC >= 8 ? 2 : 4;
  if (C.nshort > 0) {
    const int64_t rows_per_block = (int64_t)(kBlock / kWave) * (kWave / L);
    hipLaunchKernelGGL((k_spmm_rows<VC, XC, YC, RP, VEC, kRowSteps>), dim3(nblk(C.nshort, (int)rows_per_block)),
                       dim3(kBlock), 0, s, rp, C.col.as<int>(), v, vp, C.sorder.as<int>(), (int64_t)C.nshort, k, L, x, y,
                       alpha, beta);
    SPOPS_HIP_CHECK(hipGetLastError());
  }
  if (C.nslong > 0) {
    // the segments' partial sums: the caller's scratch (no allocation), else the handle's
    // buffer, grown on demand
    Acc *part = static_cast<Acc *>(ext_part);
    if (part == nullptr) {
      const size_t need = (size_t)C.nlseg * (size_t)k * sizeof(Acc);
      if (C.spmm_part.bytes < need) C.spmm_part.alloc(need);
      part = C.spmm_part.as<Acc>();
    }
    hipLaunchKernelGGL((k_spmm_seg<VC, XC, RP, VEC, kSegSteps>), dim3((unsigned)C.nlseg), dim3(kBlock), 0, s, rp,
                       C.col.as<int>(), v, vp, C.slongs.as<int>(), C.lsegs.as<int2>(), k, L, x, part);
    SPOPS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_spmm_seg_sum<Acc, YC>), dim3((unsigned)C.nslong), dim3(kBlock), 0, s, C.slongs.as<int>(),
                       C.lsegoff.as<int>(), part, k, y, alpha, beta);
    SPOPS_HIP_CHECK(hipGetLastError());
  }
}

// One vector access covers 16 bytes of X when k is a multiple of that many elements and X
// and Y are aligned for it; otherwise one element per lane.
template <int VC, int XC, int YC, typename RP>
void spmm_launch(const spops_matrix_s &A, int k, double alpha, const void *X, double beta, void *Y, hipStream_t s,
                 void *ext, const void *ev) {
  constexpr int V = 16 / (int)sizeof(typename Ty<XC>::T);
  if (k % V == 0 && aligned_to(X, 16) && aligned_to(Y, V * sizeof(typename Ty<YC>::T)))
    spmm_launch_vec<VC, XC, YC, RP, V>(A, k, alpha, X, beta, Y, s, ext, ev);
  else
    spmm_launch_vec<VC, XC, YC, RP, 1>(A, k, alpha, X, beta, Y, s, ext, ev);
}

template <int VC, int XC, typename RP>
void spmm_y(const spops_matrix_s &A, int k, double alpha, const void *X, double beta, void *Y, int yt, hipStream_t s,
            void *ext, const void *ev) {
  switch (yt) {
    case SPOPS_VALUE_F64: spmm_launch<VC, XC, SPOPS_VALUE_F64, RP>(A, k, alpha, X, beta, Y, s, ext, ev); break;
    case SPOPS_VALUE_F32: spmm_launch<VC, XC, SPOPS_VALUE_F32, RP>(A, k, alpha, X, beta, Y, s, ext, ev); break;
    default: spmm_launch<VC, XC, SPOPS_VALUE_BF16, RP>(A, k, alpha, X, beta, Y, s, ext, ev); break;
  }
}

template <int VC, typename RP>
void spmm_x(const spops_matrix_s &A, int k, double alpha, const void *X, int xt, double beta, void *Y, int yt,
            hipStream_t s, void *ext, const void *ev) {
  switch (xt) {
    case SPOPS_VALUE_F64: spmm_y<VC, SPOPS_VALUE_F64, RP>(A, k, alpha, X, beta, Y, yt, s, ext, ev); break;
    case SPOPS_VALUE_F32: spmm_y<VC, SPOPS_VALUE_F32, RP>(A, k, alpha, X, beta, Y, yt, s, ext, ev); break;
    default: spmm_y<VC, SPOPS_VALUE_BF16, RP>(A, k, alpha, X, beta, Y, yt, s, ext, ev); break;
  }
}

template <typename RP>
void spmm_v(const spops_matrix_s &A, int k, double alpha, const void *X, int xt, double beta, void *Y, int yt,
            hipStream_t s, void *ext, const void *ev) {
  switch (A.value_type) {
    case SPOPS_VALUE_F64: spmm_x<SPOPS_VALUE_F64, RP>(A, k, alpha, X, xt, beta, Y, yt, s, ext, ev); break;
    case SPOPS_VALUE_F32: spmm_x<SPOPS_VALUE_F32, RP>(A, k, alpha, X, xt, beta, Y, yt, s, ext, ev); break;
    default: spmm_x<SPOPS_VALUE_BF16, RP>(A, k, alpha, X, xt, beta, Y, yt, s, ext, ev); break;
  }
}

template <typename RP>
void build_entry_rows(const spops_matrix_s &A, hipStream_t s) {
  const auto &C = A.csr;
  C.erow.alloc((size_t)std::max<int64_t>(A.nnz, 1) * 4);
  hipLaunchKernelGGL((k_entry_rows<RP>), dim3(nblk(A.n_rows, kBlock / kWave)), dim3(kBlock), 0, s, C.rowptr.as<RP>(),
                     A.n_rows, C.erow.as<int>());
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
}

template <int GC, int XC, int OC, int VEC>
void sddmm_launch_vec(const spops_matrix_s &A, int k, const void *G, const void *X, void *out, hipStream_t s) {
  constexpr int U = 4;
  const int L = lanes_for(k, VEC);
  const int64_t per_block = (int64_t)(kBlock / kWave) * (kWave / L) * U;
  hipLaunchKernelGGL((k_sddmm<GC, XC, OC, VEC>), dim3(nblk(A.nnz, (int)per_block)), dim3(kBlock), 0, s,
                     A.csr.erow.as<int>(), A.csr.col.as<int>(), A.nnz, k, L,
                     static_cast<const typename Ty<GC>::T *>(G), static_cast<const typename Ty<XC>::T *>(X),
                     static_cast<typename Ty<OC>::T *>(out));
  SPOPS_HIP_CHECK(hipGetLastError());
}

template <int GC, int XC, int OC, typename RP, int VEC>
void sddmm_rows_launch(const spops_matrix_s &A, int k, const void *G, const void *X, void *out, hipStream_t s) {
  const auto &C = A.csr;
  const int L = lanes_for(k, VEC);
  const int64_t per_block = (int64_t)(kBlock / kWave) * (kWave / L);
  // entries per step: four with vectors of eight 2-byte elements (eight waves per SIMD), else eight
  constexpr int U = VEC >= 8 ? 4 : 8;
  hipLaunchKernelGGL((k_sddmm_rows<GC, XC, OC, RP, VEC, U>), dim3(nblk(C.nsdtasks, (int)per_block)), dim3(kBlock), 0, s,
                     C.rowptr.as<RP>(), C.col.as<int>(), C.sdtasks.as<int2>(), C.nsdtasks, k, L,
                     static_cast<const typename Ty<GC>::T *>(G), static_cast<const typename Ty<XC>::T *>(X),
                     static_cast<typename Ty<OC>::T *>(out));
  SPOPS_HIP_CHECK(hipGetLastError());
}

template <int GC, int XC, int OC>
void sddmm_launch(const spops_matrix_s &A, int k, const void *G, const void *X, void *out, hipStream_t s) {
  if (A.nnz == 0) return;
  // by rows when one lane group covers all k columns with 16-byte accesses of X (each row of
  // G loaded once, V elements per lane); else entry-parallel
  constexpr int V = 16 / (int)sizeof(typename Ty<XC>::T);
  if (k % V == 0 && k <= kWave * V && aligned_to(X, 16) && aligned_to(G, V * sizeof(typename Ty<GC>::T))) {
    if (A.csr.rowptr64) sddmm_rows_launch<GC, XC, OC, long, V>(A, k, G, X, out, s);
    else sddmm_rows_launch<GC, XC, OC, int, V>(A, k, G, X, out, s);
    return;
  }
  if (A.csr.erow.p == nullptr) {
    if (A.csr.rowptr64) build_entry_rows<long>(A, s);
    else build_entry_rows<int>(A, s);
  }
  if (k % V == 0 && aligned_to(X, 16) && aligned_to(G, V * sizeof(typename Ty<GC>::T)))
    sddmm_launch_vec<GC, XC, OC, V>(A, k, G, X, out, s);
  else
    sddmm_launch_vec<GC, XC, OC, 1>(A, k, G, X, out, s);
}

template <int GC, int XC>
void sddmm_o(const spops_matrix_s &A, int k, const void *G, const void *X, void *out, int ot, hipStream_t s) {
  switch (ot) {
    case SPOPS_VALUE_F64: sddmm_launch<GC, XC, SPOPS_VALUE_F64>(A, k, G, X, out, s); break;
    case SPOPS_VALUE_F32: sddmm_launch<GC, XC, SPOPS_VALUE_F32>(A, k, G, X, out, s); break;
    default: sddmm_launch<GC, XC, SPOPS_VALUE_BF16>(A, k, G, X, out, s); break;
  }
}

template <int GC>
void sddmm_x(const spops_matrix_s &A, int k, const void *G, const void *X, int xt, void *out, int ot, hipStream_t s) {
  switch (xt) {
    case SPOPS_VALUE_F64: sddmm_o<GC, SPOPS_VALUE_F64>(A, k, G, X, out, ot, s); break;
    case SPOPS_VALUE_F32: sddmm_o<GC, SPOPS_VALUE_F32>(A, k, G, X, out, ot, s); break;
    default: sddmm_o<GC, SPOPS_VALUE_BF16>(A, k, G, X, out, ot, s); break;
  }
}

template <int SC, int DC>
void convert_launch(const void *src, const void *perm, bool perm64, int64_t n, void *dst, hipStream_t s) {
  if (n == 0) return;
  const auto *a = static_cast<const typename Ty<SC>::T *>(src);
  auto *b = static_cast<typename Ty<DC>::T *>(dst);
  if (perm64)
    hipLaunchKernelGGL((k_convert_values<SC, DC, long>), dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, a,
                       static_cast<const long *>(perm), n, b);
  else
    hipLaunchKernelGGL((k_convert_values<SC, DC, int>), dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, a,
                       static_cast<const int *>(perm), n, b);
  SPOPS_HIP_CHECK(hipGetLastError());
}

template <int SC>
void convert_d(const void *src, const void *perm, bool perm64, int64_t n, void *dst, int dt, hipStream_t s) {
  switch (dt) {
    case SPOPS_VALUE_F64: convert_launch<SC, SPOPS_VALUE_F64>(src, perm, perm64, n, dst, s); break;
    case SPOPS_VALUE_F32: convert_launch<SC, SPOPS_VALUE_F32>(src, perm, perm64, n, dst, s); break;
    default: convert_launch<SC, SPOPS_VALUE_BF16>(src, perm, perm64, n, dst, s); break;
  }
}

// The plan of y[i] = x[g[i]] (g a permutation of 0..n-1; see PermPlan).  Source j goes to
// destination d = inv[j], tile d / S.  Pass 1 orders each source tile by destination tile
// (loc1, stable); jof[a] is the source at staging address a.  Pass 2's list for destination tile
// b walks the staging addresses in order and keeps those whose source goes to b (src2), with
// their places in the tile (loc2).
void build_perm_plan(const std::vector<int32_t> &g, int64_t S, PermPlan &P) {
  const int64_t n = (int64_t)g.size();
  const int64_t nt = (n + S - 1) / S;
  std::vector<int32_t> inv((size_t)n), jof((size_t)n), src2((size_t)n);
  std::vector<uint16_t> loc1((size_t)n), loc2((size_t)n);
  for (int64_t i = 0; i < n; ++i) inv[(size_t)g[(size_t)i]] = (int32_t)i;
  std::vector<int64_t> cnt((size_t)nt + 1);
  for (int64_t j0 = 0; j0 < n; j0 += S) {
    const int64_t j1 = std::min(n, j0 + S);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t j = j0; j < j1; ++j) ++cnt[(size_t)(inv[(size_t)j] / S) + 1];
    for (int64_t b = 0; b < nt; ++b) cnt[(size_t)b + 1] += cnt[(size_t)b];
    for (int64_t j = j0; j < j1; ++j) {
      const int64_t r = cnt[(size_t)(inv[(size_t)j] / S)]++;   // place in the tile ordered by destination tile
      loc1[(size_t)j] = (uint16_t)r;
      jof[(size_t)(j0 + r)] = (int32_t)j;
    }
  }
  std::vector<int64_t> cur((size_t)nt);
  for (int64_t b = 0; b < nt; ++b) cur[(size_t)b] = b * S;
  for (int64_t a = 0; a < n; ++a) {
    const int64_t d = inv[(size_t)jof[(size_t)a]];
    const int64_t r = cur[(size_t)(d / S)]++;
    src2[(size_t)r] = (int32_t)((a / S) * (S + kPermPad) + a % S);
    loc2[(size_t)r] = (uint16_t)(d % S);
  }
  P.loc1.upload(loc1);
  P.src2.upload(src2);
  P.loc2.upload(loc2);
  P.n = n;
  P.tile = S;
}

// tile of the two-pass permutation: 128 KiB of LDS (16384 elements of 8 bytes, 32768 of 4 or 2)
inline int64_t perm_tile(size_t esize) { return esize == 8 ? kPermTile : 2 * kPermTile; }

// Both permutations of a transposed handle (its values from the source order, and its entry
// order back to the source's), built from tperm on first use.
void ensure_perm_plans(const spops_matrix_s &AT) {
  const auto &C = AT.csr;
  if (C.pin.n == AT.nnz && C.pout.n == AT.nnz) return;
  const int64_t S = perm_tile(value_bytes(AT.value_type));
  std::vector<int32_t> tp((size_t)AT.nnz), ti((size_t)AT.nnz);
  if (AT.nnz) SPOPS_HIP_CHECK(hipMemcpy(tp.data(), C.tperm.p, (size_t)AT.nnz * 4, hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < AT.nnz; ++e) ti[(size_t)tp[(size_t)e]] = (int32_t)e;
  build_perm_plan(tp, S, C.pin);
  build_perm_plan(ti, S, C.pout);
}

void perm_apply(const PermPlan &P, const void *x, void *y, size_t esize, DevBuf &stage, hipStream_t s) {
  const int64_t n = P.n;
  if (n <= 0) return;
  SPOPS_REQUIRE(P.tile == perm_tile(esize), SPOPS_ERROR_INVALID_VALUE, "permutation plan built for another element size");
  const unsigned tiles = (unsigned)((n + P.tile - 1) / P.tile);
  const size_t need = (size_t)tiles * (size_t)(P.tile + kPermPad) * esize;
  SPOPS_REQUIRE(need < kMaxBufBytes, SPOPS_ERROR_LIMIT, "two-pass permutation: staging of >= 2 GiB");
  if (stage.bytes < need) stage.alloc(need);
  const auto *l1 = P.loc1.as<uint16_t>();
  const auto *l2 = P.loc2.as<uint16_t>();
  const int *s2 = P.src2.as<int>();
  auto run = [&](auto tag, auto ts) {
    using T = decltype(tag);
    constexpr int TS = decltype(ts)::value;
    hipLaunchKernelGGL((k_perm_tiles<T, TS>), dim3(tiles), dim3(kPermThreads), 0, s, static_cast<const T *>(x), l1, n,
                       stage.as<T>());
    SPOPS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_perm_buckets<T, TS>), dim3(tiles), dim3(kPermThreads), 0, s, stage.as<T>(),
                       (int64_t)(need / esize), s2, l2, n, static_cast<T *>(y));
    SPOPS_HIP_CHECK(hipGetLastError());
  };
  if (esize == 8) run(uint64_t{}, std::integral_constant<int, kPermTile>{});
  else if (esize == 4) run(uint32_t{}, std::integral_constant<int, 2 * kPermTile>{});
  else run(uint16_t{}, std::integral_constant<int, 2 * kPermTile>{});
}

// The fused backward on a transposed handle (k_spmm_sddmm_rows / _seg, then k_spmm_seg_sum for
// the long rows' dX), the dot products in AT's order into the handle's buffer, then gathered
// into the source matrix's order through the inverse permutation (built on first use); the
// caller has checked that one lane group covers k with 16-byte G loads.
template <int VC, int GC, int XC>
void spmm_sddmm_launch(const spops_matrix_s &AT, int k, const void *G, const void *X, void *dX, void *dvals,
                       hipStream_t s) {
  constexpr int V = 16 / (int)sizeof(typename Ty<GC>::T);
  using Acc = AccT<VC, GC>;
  using VT = typename Ty<VC>::T;
  const auto &C = AT.csr;
  const int L = lanes_for(k, V);
  const int *rp = C.rowptr.as<int>();
  const auto *v = C.val.as<VT>();
  const auto *g = static_cast<const typename Ty<GC>::T *>(G);
  const auto *x = static_cast<const typename Ty<XC>::T *>(X);
  auto *dx = static_cast<typename Ty<XC>::T *>(dX);
  if (AT.nnz > 0) {
    ensure_perm_plans(AT);
    if (C.tdv.bytes < (size_t)AT.nnz * sizeof(VT)) C.tdv.alloc((size_t)AT.nnz * sizeof(VT));
  }
  VT *dT = C.tdv.as<VT>();
  if (C.nshort > 0) {
    const int64_t rows_per_block = (int64_t)(kBlock / kWave) * (kWave / L);
    // four entries per step (FP32 G: 59 VGPRs, eight waves per SIMD; eight entries: 99 VGPRs, four
    // waves, 610 against 536 us at R-MAT 20; BF16 G and X with packed dot products: 63 VGPRs, eight
    // waves, 378 us, against 83 VGPRs, five waves, 461 us with the products in FP32)
    hipLaunchKernelGGL((k_spmm_sddmm_rows<VC, GC, XC, int, V, 4>), dim3(nblk(C.nshort, (int)rows_per_block)),
                       dim3(kBlock), 0, s, rp, C.col.as<int>(), v, C.sorder.as<int>(), (int64_t)C.nshort, k, L, g, x, dx,
                       dT);
    SPOPS_HIP_CHECK(hipGetLastError());
  }
  if (C.nslong > 0) {
    const size_t need = (size_t)C.nlseg * (size_t)k * sizeof(Acc);
    if (C.spmm_part.bytes < need) C.spmm_part.alloc(need);
    Acc *part = C.spmm_part.as<Acc>();
    // four entries per step, but two with BF16 G and FP32 / FP64 X (56 VGPRs, eight waves per SIMD;
    // four: 77 VGPRs, six waves, 418 against 371 us at R-MAT 20, k = 64); BF16 G and X with packed
    // dot products: 57 VGPRs at four, 301 us
    hipLaunchKernelGGL((k_spmm_sddmm_seg<VC, GC, XC, int, V, (V >= 8 && XC != SPOPS_VALUE_BF16 ? 2 : 4)>), dim3((unsigned)C.nlseg), dim3(kBlock), 0, s, rp,
                       C.col.as<int>(), v, C.slongs.as<int>(), C.lsegs.as<int2>(), k, L, g, x, part, dT);
    SPOPS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_spmm_seg_sum<Acc, XC>), dim3((unsigned)C.nslong), dim3(kBlock), 0, s, C.slongs.as<int>(),
                       C.lsegoff.as<int>(), part, k, dx, 1.0, 0.0);
    SPOPS_HIP_CHECK(hipGetLastError());
  }
  if (AT.nnz > 0) perm_apply(C.pout, dT, dvals, sizeof(VT), C.pstage, s);
}

template <int VC, int GC>
void spmm_sddmm_x(const spops_matrix_s &AT, int k, const void *G, const void *X, int xt, void *dX, void *dvals,
                  hipStream_t s) {
  switch (xt) {
    case SPOPS_VALUE_F64: spmm_sddmm_launch<VC, GC, SPOPS_VALUE_F64>(AT, k, G, X, dX, dvals, s); break;
    case SPOPS_VALUE_F32: spmm_sddmm_launch<VC, GC, SPOPS_VALUE_F32>(AT, k, G, X, dX, dvals, s); break;
    default: spmm_sddmm_launch<VC, GC, SPOPS_VALUE_BF16>(AT, k, G, X, dX, dvals, s); break;
  }
}

template <int VC>
void spmm_sddmm_g(const spops_matrix_s &AT, int k, const void *G, int gt, const void *X, int xt, void *dX,
                  void *dvals, hipStream_t s) {
  switch (gt) {
    case SPOPS_VALUE_F64: spmm_sddmm_x<VC, SPOPS_VALUE_F64>(AT, k, G, X, xt, dX, dvals, s); break;
    case SPOPS_VALUE_F32: spmm_sddmm_x<VC, SPOPS_VALUE_F32>(AT, k, G, X, xt, dX, dvals, s); break;
    default: spmm_sddmm_x<VC, SPOPS_VALUE_BF16>(AT, k, G, X, xt, dX, dvals, s); break;
  }
}

void convert_values(const void *src, int st, const void *perm, bool perm64, int64_t n, void *dst, int dt,
                    hipStream_t s) {
  switch (st) {
    case SPOPS_VALUE_F64: convert_d<SPOPS_VALUE_F64>(src, perm, perm64, n, dst, dt, s); break;
    case SPOPS_VALUE_F32: convert_d<SPOPS_VALUE_F32>(src, perm, perm64, n, dst, dt, s); break;
    default: convert_d<SPOPS_VALUE_BF16>(src, perm, perm64, n, dst, dt, s); break;
  }
}

}  // namespace

void launch_spmm(const spops_matrix_s &A, int64_t k, double alpha, const void *X, int xt, double beta, void *Y, int yt,
                 hipStream_t s, void *ext_part, const void *ext_val) {
  require_csr(&A);
  require_type(xt, "X");
  require_type(yt, "Y");
  SPOPS_REQUIRE(k > 0 && k < (1 << 20), SPOPS_ERROR_INVALID_VALUE, "k must be in [1, 2^20)");
  SPOPS_REQUIRE((uint64_t)A.n_cols * (uint64_t)k < ((uint64_t)1 << 40) &&
                    (uint64_t)A.n_rows * (uint64_t)k < ((uint64_t)1 << 40),
                SPOPS_ERROR_LIMIT, "n * k too large");
  if (A.csr.rowptr64) spmm_v<long>(A, (int)k, alpha, X, xt, beta, Y, yt, s, ext_part, ext_val);
  else spmm_v<int>(A, (int)k, alpha, X, xt, beta, Y, yt, s, ext_part, ext_val);
}

}  // namespace spops

// Scratch of spops_spmm_typed_noalloc: the partial sums of the long rows' segments (one row
// of k sums per segment of kSpmmSegment entries, FP64 or FP32).  Each GPU holds its own copy
// for the life of the process; every call on that GPU shares it.
constexpr size_t kSpmmScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kSpmmScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (the array's name is a host-side
// stand-in there), looked up on the first call on each device and kept.
void *spmm_scratch_address() {
  int dev = 0;
  SPOPS_HIP_CHECK(hipGetDevice(&dev));
  SPOPS_REQUIRE(dev >= 0 && dev < kMaxDevices, SPOPS_ERROR_LIMIT, "device index %d out of range", dev);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    SPOPS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SPOPS_REQUIRE(p != nullptr, SPOPS_ERROR_HIP, "no device address for the SpMM scratch");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

using namespace spops;

extern "C" int spops_spmm(spops_matrix A, int64_t k, double alpha, const void *d_X, double beta, void *d_Y,
                          void *stream) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    const int t = A->value_type == SPOPS_VALUE_F64 ? SPOPS_VALUE_F64 : SPOPS_VALUE_F32;
    launch_spmm(*A, k, alpha, d_X, t, beta, d_Y, t, static_cast<hipStream_t>(stream), nullptr, nullptr);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmm_typed(spops_matrix A, int64_t k, double alpha, const void *d_X, int x_type, double beta,
                                void *d_Y, int y_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    launch_spmm(*A, k, alpha, d_X, x_type, beta, d_Y, y_type, static_cast<hipStream_t>(stream), nullptr, nullptr);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmm_values(spops_matrix A, const void *d_values, int64_t k, double alpha, const void *d_X,
                                 int x_type, double beta, void *d_Y, int y_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE(d_values != nullptr || A->nnz == 0, SPOPS_ERROR_INVALID_VALUE, "null values");
    SPOPS_REQUIRE(!A->csr.tperm64, SPOPS_ERROR_LIMIT,
                  "spops_spmm_values: a transposed handle of >= 2^31 entries (use spops_csr_set_values)");
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    launch_spmm(*A, k, alpha, d_X, x_type, beta, d_Y, y_type, static_cast<hipStream_t>(stream), nullptr,
                A->nnz > 0 ? d_values : nullptr);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmm_typed_noalloc(spops_matrix A, int64_t k, double alpha, const void *d_X, int x_type,
                                        double beta, void *d_Y, int y_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    require_type(x_type, "X");
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    SPOPS_REQUIRE(k > 0 && k < (1 << 20), SPOPS_ERROR_INVALID_VALUE, "k must be in [1, 2^20)");
    const size_t acc = (A->value_type == SPOPS_VALUE_F64 || x_type == SPOPS_VALUE_F64) ? 8 : 4;
    const size_t need = (size_t)A->csr.nlseg * (size_t)k * acc;
    SPOPS_REQUIRE(need <= kSpmmScratchBytes, SPOPS_ERROR_LIMIT,
                  "spops_spmm_typed_noalloc: %zu bytes of segment partial sums (%d segments x k=%lld) exceed its "
                  "static scratch of %zu bytes; use spops_spmm_typed",
                  need, A->csr.nlseg, (long long)k, kSpmmScratchBytes);
    void *part = A->csr.nlseg > 0 ? spmm_scratch_address() : nullptr;
    launch_spmm(*A, k, alpha, d_X, x_type, beta, d_Y, y_type, static_cast<hipStream_t>(stream),
                part, nullptr);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_sddmm(spops_matrix A, int64_t k, const void *d_G, int g_type, const void *d_X, int x_type,
                           void *d_out, int out_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    require_type(g_type, "G");
    require_type(x_type, "X");
    require_type(out_type, "out");
    SPOPS_REQUIRE(k > 0 && k < (1 << 20), SPOPS_ERROR_INVALID_VALUE, "k must be in [1, 2^20)");
    SPOPS_REQUIRE(A->nnz == 0 || (d_G != nullptr && d_X != nullptr && d_out != nullptr), SPOPS_ERROR_INVALID_VALUE,
                  "null G, X or out");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (g_type) {
      case SPOPS_VALUE_F64: sddmm_x<SPOPS_VALUE_F64>(*A, (int)k, d_G, d_X, x_type, d_out, out_type, s); break;
      case SPOPS_VALUE_F32: sddmm_x<SPOPS_VALUE_F32>(*A, (int)k, d_G, d_X, x_type, d_out, out_type, s); break;
      default: sddmm_x<SPOPS_VALUE_BF16>(*A, (int)k, d_G, d_X, x_type, d_out, out_type, s); break;
    }
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmm_sddmm_transposed(spops_matrix AT, int64_t k, const void *d_G, int g_type, const void *d_X,
                                          int x_type, void *d_dX, void *d_dvals, void *stream) {
  return guarded([&]() {
    require_csr(AT);
    require_type(g_type, "G");
    require_type(x_type, "X");
    SPOPS_REQUIRE(AT->csr.tperm.p != nullptr || AT->nnz == 0, SPOPS_ERROR_INVALID_VALUE,
                  "spops_spmm_sddmm_transposed: AT must come from spops_csr_transpose");
    SPOPS_REQUIRE(!AT->csr.tperm64, SPOPS_ERROR_NOT_SUPPORTED,
                  "spops_spmm_sddmm_transposed: a transposed handle of >= 2^31 entries");
    SPOPS_REQUIRE(k > 0 && k < (1 << 20), SPOPS_ERROR_INVALID_VALUE, "k must be in [1, 2^20)");
    SPOPS_REQUIRE((d_G != nullptr || AT->n_cols == 0) && (d_X != nullptr && d_dX != nullptr || AT->n_rows == 0) &&
                      (d_dvals != nullptr || AT->nnz == 0),
                  SPOPS_ERROR_INVALID_VALUE, "null G, X, dX or dvals");
    // one lane group over all k columns with 16-byte loads of G and aligned rows of X and dX
    const size_t gs = value_bytes(g_type), xs = value_bytes(x_type);
    const int64_t V = 16 / (int64_t)gs;
    const size_t xa = (size_t)V * xs;
    SPOPS_REQUIRE(k % V == 0 && k <= kWave * V && ((uintptr_t)d_G % 16) == 0 && ((uintptr_t)d_X % xa) == 0 &&
                      ((uintptr_t)d_dX % xa) == 0,
                  SPOPS_ERROR_NOT_SUPPORTED,
                  "spops_spmm_sddmm_transposed: k = %lld needs to be a multiple of %lld and at most %lld, G 16-byte "
                  "aligned and X, dX %zu-byte aligned (use spops_spmm_typed and spops_sddmm)",
                  (long long)k, (long long)V, (long long)(kWave * V), xa);
    // the two-pass permutation's staging array is addressed with 32-bit offsets
    const int64_t S = perm_tile(value_bytes(AT->value_type));
    const int64_t tiles = (AT->nnz + S - 1) / S;
    SPOPS_REQUIRE((uint64_t)tiles * (uint64_t)(S + kPermPad) * value_bytes(AT->value_type) < kMaxBufBytes,
                  SPOPS_ERROR_NOT_SUPPORTED,
                  "spops_spmm_sddmm_transposed: %lld entries need a staging array of >= 2 GiB (use spops_spmm_typed and "
                  "spops_sddmm)", (long long)AT->nnz);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (AT->value_type) {
      case SPOPS_VALUE_F64: spmm_sddmm_g<SPOPS_VALUE_F64>(*AT, (int)k, d_G, g_type, d_X, x_type, d_dX, d_dvals, s); break;
      case SPOPS_VALUE_F32: spmm_sddmm_g<SPOPS_VALUE_F32>(*AT, (int)k, d_G, g_type, d_X, x_type, d_dX, d_dvals, s); break;
      default: spmm_sddmm_g<SPOPS_VALUE_BF16>(*AT, (int)k, d_G, g_type, d_X, x_type, d_dX, d_dvals, s); break;
    }
    return SPOPS_SUCCESS;
  });
}

namespace spops {
void refresh_sweep_values(spops_matrix_s &A, hipStream_t s) {
  const int64_t slots = (int64_t)A.sweep.ngroups * kSweepGroup;
  if (A.sweep.H == 0 || slots == 0 || A.sweep.src.p == nullptr) return;
  const int32_t *src = A.sweep.src.as<int32_t>();
  switch (value_bytes(A.value_type)) {
    case 8:
      hipLaunchKernelGGL(k_sweep_refresh<uint64_t>, dim3(nblk(slots, kBlock)), dim3(kBlock), 0, s, src, slots,
                         A.csr.val.as<uint64_t>(), A.sweep.val.as<uint64_t>());
      break;
    case 4:
      hipLaunchKernelGGL(k_sweep_refresh<uint32_t>, dim3(nblk(slots, kBlock)), dim3(kBlock), 0, s, src, slots,
                         A.csr.val.as<uint32_t>(), A.sweep.val.as<uint32_t>());
      break;
    default:
      hipLaunchKernelGGL(k_sweep_refresh<uint16_t>, dim3(nblk(slots, kBlock)), dim3(kBlock), 0, s, src, slots,
                         A.csr.val.as<uint16_t>(), A.sweep.val.as<uint16_t>());
  }
  SPOPS_HIP_CHECK(hipGetLastError());
}
}  // namespace spops

extern "C" int spops_csr_set_values(spops_matrix A, const void *d_values, int values_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    require_type(values_type, "values");
    SPOPS_REQUIRE(A->nnz == 0 || d_values != nullptr, SPOPS_ERROR_INVALID_VALUE, "null values");
    const bool perm = A->csr.tperm.p != nullptr;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (perm && values_type == A->value_type && A->csr.pin.n == A->nnz)
      // a transposed handle that ran the fused backward keeps the two-pass form of its permutation
      perm_apply(A->csr.pin, d_values, A->csr.val.p, value_bytes(A->value_type), A->csr.pstage, s);
    else
      convert_values(d_values, values_type, perm ? A->csr.tperm.p : nullptr, A->csr.tperm64, A->nnz, A->csr.val.p,
                     A->value_type, s);
    // the swept SpMV layout's copy follows on the same stream, so any later SpMV ordered after
    // this call (or captured after it) reads the new values; handles made for SpMM / SDDMM only
    // (spops_csr_create_layout with spmv_layout 1) have no such copy and pay nothing here
    if (A->sweep.H > 0 && A->slots > 0) refresh_sweep_values(*A, s);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_csr_transpose(spops_matrix A, spops_matrix *AT) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE(AT != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output handle");
    SPOPS_REQUIRE(A->csr.tperm.p == nullptr, SPOPS_ERROR_NOT_SUPPORTED, "A is itself a transposed view");
    const int64_t n = A->n_rows, m = A->n_cols, nnz = A->nnz;
    // structure to the host (the transpose is built once per pattern)
    std::vector<int64_t> rp(n + 1);
    if (A->csr.rowptr64) {
      SPOPS_HIP_CHECK(hipMemcpy(rp.data(), A->csr.rowptr.p, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<int32_t> r32(n + 1);
      SPOPS_HIP_CHECK(hipMemcpy(r32.data(), A->csr.rowptr.p, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
      std::copy(r32.begin(), r32.end(), rp.begin());
    }
    std::vector<int32_t> ci(nnz);
    if (nnz) SPOPS_HIP_CHECK(hipMemcpy(ci.data(), A->csr.col.p, (size_t)nnz * 4, hipMemcpyDeviceToHost));
    // counting sort by column, stable in the row (csr_transpose's order, sparse_conversion.h:2067-2108)
    std::vector<int64_t> trp(m + 1, 0), perm(nnz);
    std::vector<int32_t> tcol(nnz);
    for (int64_t j = 0; j < nnz; ++j) ++trp[ci[j] + 1];
    for (int64_t c = 0; c < m; ++c) trp[c + 1] += trp[c];
    std::vector<int64_t> fill(trp.begin(), trp.end() - 1);
    for (int64_t r = 0; r < n; ++r)
      for (int64_t j = rp[r]; j < rp[r + 1]; ++j) {
        const int64_t at = fill[ci[j]]++;
        tcol[at] = (int32_t)r;
        perm[at] = j;
      }
    auto T = std::make_unique<spops_matrix_s>();
    T->format = SPOPS_FORMAT_CSR;
    T->value_type = A->value_type;
    T->index_type = A->index_type;
    T->n_rows = m;
    T->n_cols = n;
    T->nnz = nnz;
    auto &P = T->csr;
    P.rowptr64 = nnz >= (int64_t)INT32_MAX;
    if (P.rowptr64) P.rowptr.upload(trp);
    else P.rowptr.upload(std::vector<int32_t>(trp.begin(), trp.end()));
    P.col.upload(tcol);
    P.tperm64 = P.rowptr64;
    if (P.tperm64) P.tperm.upload(perm);
    else P.tperm.upload(std::vector<int32_t>(perm.begin(), perm.end()));
    P.val.alloc((size_t)std::max<int64_t>(nnz, 1) * value_bytes(A->value_type));
    // values follow A's through the permutation
    convert_values(A->csr.val.p, A->value_type, P.tperm.p, P.tperm64, nnz, P.val.p, A->value_type, nullptr);
    SPOPS_HIP_CHECK(hipDeviceSynchronize());
    upload_csr_plan(P, trp, m);
    T->slots = nnz;
    T->dev_bytes = P.rowptr.bytes + P.col.bytes + P.val.bytes + P.tasks.bytes + P.longs.bytes + P.tperm.bytes;
    T->algo_bytes = spmv_algo_bytes(*T, A->index_type == SPOPS_INDEX_I32 ? 4 : 8);
    *AT = T.release();
    return SPOPS_SUCCESS;
  });
}
