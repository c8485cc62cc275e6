// SpMM, SDDMM, transpose and value updates on CSR matrices: the multi-RHS and the
// differentiable forms of SpOps_csr (src/spmv_csr.cpp:165-185 applied column by column).
//
//   Y = alpha * A * X + beta * Y        X: n_cols x k, Y: n_rows x k, both row-major
//   out[e] = sum_c G[row_e, c] * X[col_e, c]      (SDDMM on A's pattern: d(loss)/d(values))
//   A^T with the entry permutation kept, so that A^T's values follow A's (dX = A^T * G)
//
// X, Y, G in FP64, FP32 or BF16; accumulation in FP64 when the matrix or X is FP64, else
// FP32.  Kernels:
//   k_spmm_rows : one row per lane group of GS lanes (GS = next power of two >= k, capped at
//                 64; 64 / GS rows per wave), lanes over the k columns, eight row entries
//                 in flight (their X rows are independent loads).  Rows longer than
//                 kCsrLongRow entries are left to
//   k_spmm_long : one workgroup per long row, the 4 waves split its entries, LDS reduction.
//   k_sddmm     : one wave per row, lanes over the k columns; per entry a wave reduction.
#include <algorithm>
#include <memory>
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
  // float -> bf16, round to nearest even (the device twin of float_to_bf16_rne)
  static __device__ __forceinline__ uint16_t st(double d) {
    const float f = (float)d;
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
};

template <int VC, int XC>
using AccT = typename std::conditional<VC == SPOPS_VALUE_F64 || XC == SPOPS_VALUE_F64, double, float>::type;

template <int VC, int XC, int YC, typename RP, int GS>
__global__ __launch_bounds__(kBlock) void k_spmm_rows(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                      const typename Ty<VC>::T *__restrict__ val, int64_t n_rows, int k,
                                                      const typename Ty<XC>::T *__restrict__ X,
                                                      typename Ty<YC>::T *__restrict__ Y, double alpha, double beta) {
  using Acc = AccT<VC, XC>;
  constexpr int R = kWave / GS, U = 8;
  const int lane = lane_id(), g = lane / GS, c0 = lane % GS;
  const int64_t row = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * R + g;
  if (row >= n_rows) return;                        // lane groups past the last row
  const int64_t p0 = (int64_t)rowptr[row], p1 = (int64_t)rowptr[row + 1];
  if (p1 - p0 > kCsrLongRow) return;                // a long row: k_spmm_long writes it
  for (int cb = 0; cb < k; cb += GS) {
    const int c = cb + c0;
    const bool cv = c < k;
    Acc acc = 0;
    for (int64_t p = p0; p < p1; p += U) {
      int cc[U];
      Acc vv[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = p + u < p1;
        cc[u] = ok ? col[p + u] : 0;
        vv[u] = ok ? (Acc)Ty<VC>::ld(val[p + u]) : Acc(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        xv[u] = (cv && p + u < p1) ? (Acc)Ty<XC>::ld(X[(int64_t)cc[u] * k + c]) : Acc(0);
#pragma unroll
      for (int u = 0; u < U; ++u) acc = vv[u] * xv[u] + acc;
    }
    if (cv) {
      double o = alpha * (double)acc;
      if (beta != 0.0) o += beta * Ty<YC>::ld(Y[row * k + c]);
      Y[row * k + c] = Ty<YC>::st(o);
    }
  }
}

template <int VC, int XC, int YC, typename RP>
__global__ __launch_bounds__(kBlock) void k_spmm_long(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                      const typename Ty<VC>::T *__restrict__ val,
                                                      const int *__restrict__ long_rows, int k,
                                                      const typename Ty<XC>::T *__restrict__ X,
                                                      typename Ty<YC>::T *__restrict__ Y, double alpha, double beta) {
  using Acc = AccT<VC, XC>;
  constexpr int W = kBlock / kWave, U = 8;
  __shared__ Acc red[W][kWave];
  const int64_t row = long_rows[blockIdx.x];
  const int64_t p0 = (int64_t)rowptr[row], p1 = (int64_t)rowptr[row + 1];
  const int lane = lane_id(), w = wave_id();
  for (int cb = 0; cb < k; cb += kWave) {
    const int c = cb + lane;
    const bool cv = c < k;
    Acc acc = 0;
    for (int64_t p = p0 + (int64_t)w * U; p < p1; p += (int64_t)W * U) {
      int cc[U];
      Acc vv[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = p + u < p1;
        cc[u] = ok ? col[p + u] : 0;
        vv[u] = ok ? (Acc)Ty<VC>::ld(val[p + u]) : Acc(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        xv[u] = (cv && p + u < p1) ? (Acc)Ty<XC>::ld(X[(int64_t)cc[u] * k + c]) : Acc(0);
#pragma unroll
      for (int u = 0; u < U; ++u) acc = vv[u] * xv[u] + acc;
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && cv) {
      Acc s = 0;
      for (int q = 0; q < W; ++q) s += red[q][lane];
      double o = alpha * (double)s;
      if (beta != 0.0) o += beta * Ty<YC>::ld(Y[row * k + c]);
      Y[row * k + c] = Ty<YC>::st(o);
    }
    __syncthreads();
  }
}

// out[e] = sum_c G[row_e, c] * X[col_e, c]; one wave per row
template <int GC, int XC, int OC, typename RP>
__global__ __launch_bounds__(kBlock) void k_sddmm(const RP *__restrict__ rowptr, const int *__restrict__ col,
                                                  int64_t n_rows, int k, const typename Ty<GC>::T *__restrict__ G,
                                                  const typename Ty<XC>::T *__restrict__ X,
                                                  typename Ty<OC>::T *__restrict__ out) {
  using Acc = AccT<GC, XC>;
  constexpr int U = 8;
  const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (row >= n_rows) return;
  const int lane = lane_id();
  const int64_t p0 = (int64_t)rowptr[row], p1 = (int64_t)rowptr[row + 1];
  for (int64_t p = p0; p < p1; p += U) {
    int cc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) cc[u] = p + u < p1 ? col[p + u] : 0;
    Acc part[U];
#pragma unroll
    for (int u = 0; u < U; ++u) part[u] = 0;
    for (int cb = 0; cb < k; cb += kWave) {
      const int c = cb + lane;
      const bool cv = c < k;
      const Acc gv = cv ? (Acc)Ty<GC>::ld(G[row * k + c]) : Acc(0);
#pragma unroll
      for (int u = 0; u < U; ++u)
        part[u] += (cv && p + u < p1) ? gv * (Acc)Ty<XC>::ld(X[(int64_t)cc[u] * k + c]) : Acc(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) part[u] += __shfl_xor(part[u], m, kWave);
    }
    if (lane < U && p + lane < p1) {
      Acc v = part[0];
#pragma unroll
      for (int u = 1; u < U; ++u) v = lane == u ? part[u] : v;
      out[p + lane] = Ty<OC>::st((double)v);
    }
  }
}

template <int SC, int DC, typename IDX>
__global__ __launch_bounds__(kBlock) void k_convert_values(const typename Ty<SC>::T *__restrict__ src,
                                                           const IDX *__restrict__ perm, int64_t n,
                                                           typename Ty<DC>::T *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dst[i] = Ty<DC>::st(Ty<SC>::ld(src[perm ? (int64_t)perm[i] : i]));
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

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
template <int VC, int XC, int YC, typename RP>
void spmm_launch(const spops_matrix_s &A, int k, double alpha, const void *X, double beta, void *Y, hipStream_t s) {
  const auto &C = A.csr;
  const RP *rp = C.rowptr.as<RP>();
  const auto *v = C.val.as<typename Ty<VC>::T>();
  const auto *x = static_cast<const typename Ty<XC>::T *>(X);
  auto *y = static_cast<typename Ty<YC>::T *>(Y);
  const int64_t n = A.n_rows;
  if (n > 0) {
    const int GS = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
    const int64_t rows_per_block = (int64_t)(kBlock / kWave) * (kWave / GS);
    const unsigned grid = nblk(n, (int)rows_per_block);
#define SPOPS_SPMM_ROWS(G)                                                                                    \
  hipLaunchKernelGGL((k_spmm_rows<VC, XC, YC, RP, G>), dim3(grid), dim3(kBlock), 0, s, rp, C.col.as<int>(), v, n, k, \
                     x, y, alpha, beta)
    switch (GS) {
      case 4: SPOPS_SPMM_ROWS(4); break;
      case 8: SPOPS_SPMM_ROWS(8); break;
      case 16: SPOPS_SPMM_ROWS(16); break;
      case 32: SPOPS_SPMM_ROWS(32); break;
      default: SPOPS_SPMM_ROWS(64); break;
    }
#undef SPOPS_SPMM_ROWS
    SPOPS_HIP_CHECK(hipGetLastError());
  }
  if (C.nlong > 0) {
    hipLaunchKernelGGL((k_spmm_long<VC, XC, YC, RP>), dim3((unsigned)C.nlong), dim3(kBlock), 0, s, rp,
                       C.col.as<int>(), v, C.longs.as<int>(), k, x, y, alpha, beta);
    SPOPS_HIP_CHECK(hipGetLastError());
  }
}

template <int VC, int XC, typename RP>
void spmm_y(const spops_matrix_s &A, int k, double alpha, const void *X, double beta, void *Y, int yt, hipStream_t s) {
  switch (yt) {
    case SPOPS_VALUE_F64: spmm_launch<VC, XC, SPOPS_VALUE_F64, RP>(A, k, alpha, X, beta, Y, s); break;
    case SPOPS_VALUE_F32: spmm_launch<VC, XC, SPOPS_VALUE_F32, RP>(A, k, alpha, X, beta, Y, s); break;
    default: spmm_launch<VC, XC, SPOPS_VALUE_BF16, RP>(A, k, alpha, X, beta, Y, s); break;
  }
}

template <int VC, typename RP>
void spmm_x(const spops_matrix_s &A, int k, double alpha, const void *X, int xt, double beta, void *Y, int yt,
            hipStream_t s) {
  switch (xt) {
    case SPOPS_VALUE_F64: spmm_y<VC, SPOPS_VALUE_F64, RP>(A, k, alpha, X, beta, Y, yt, s); break;
    case SPOPS_VALUE_F32: spmm_y<VC, SPOPS_VALUE_F32, RP>(A, k, alpha, X, beta, Y, yt, s); break;
    default: spmm_y<VC, SPOPS_VALUE_BF16, RP>(A, k, alpha, X, beta, Y, yt, s); break;
  }
}

template <typename RP>
void spmm_v(const spops_matrix_s &A, int k, double alpha, const void *X, int xt, double beta, void *Y, int yt,
            hipStream_t s) {
  switch (A.value_type) {
    case SPOPS_VALUE_F64: spmm_x<SPOPS_VALUE_F64, RP>(A, k, alpha, X, xt, beta, Y, yt, s); break;
    case SPOPS_VALUE_F32: spmm_x<SPOPS_VALUE_F32, RP>(A, k, alpha, X, xt, beta, Y, yt, s); break;
    default: spmm_x<SPOPS_VALUE_BF16, RP>(A, k, alpha, X, xt, beta, Y, yt, s); break;
  }
}

template <int GC, int XC, int OC>
void sddmm_launch(const spops_matrix_s &A, int k, const void *G, const void *X, void *out, hipStream_t s) {
  const int64_t n = A.n_rows;
  if (n == 0) return;
  const unsigned grid = nblk(n, kBlock / kWave);
  const auto *g = static_cast<const typename Ty<GC>::T *>(G);
  const auto *x = static_cast<const typename Ty<XC>::T *>(X);
  auto *o = static_cast<typename Ty<OC>::T *>(out);
  if (A.csr.rowptr64)
    hipLaunchKernelGGL((k_sddmm<GC, XC, OC, long>), dim3(grid), dim3(kBlock), 0, s, A.csr.rowptr.as<long>(),
                       A.csr.col.as<int>(), n, k, g, x, o);
  else
    hipLaunchKernelGGL((k_sddmm<GC, XC, OC, int>), dim3(grid), dim3(kBlock), 0, s, A.csr.rowptr.as<int>(),
                       A.csr.col.as<int>(), n, k, g, x, o);
  SPOPS_HIP_CHECK(hipGetLastError());
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
                 hipStream_t s) {
  require_csr(&A);
  require_type(xt, "X");
  require_type(yt, "Y");
  SPOPS_REQUIRE(k > 0 && k < (1 << 20), SPOPS_ERROR_INVALID_VALUE, "k must be in [1, 2^20)");
  SPOPS_REQUIRE((uint64_t)A.n_cols * (uint64_t)k < ((uint64_t)1 << 40) &&
                    (uint64_t)A.n_rows * (uint64_t)k < ((uint64_t)1 << 40),
                SPOPS_ERROR_LIMIT, "n * k too large");
  if (A.csr.rowptr64) spmm_v<long>(A, (int)k, alpha, X, xt, beta, Y, yt, s);
  else spmm_v<int>(A, (int)k, alpha, X, xt, beta, Y, yt, s);
}

}  // namespace spops

using namespace spops;

extern "C" int spops_spmm(spops_matrix A, int64_t k, double alpha, const void *d_X, double beta, void *d_Y,
                          void *stream) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    const int t = A->value_type == SPOPS_VALUE_F64 ? SPOPS_VALUE_F64 : SPOPS_VALUE_F32;
    launch_spmm(*A, k, alpha, d_X, t, beta, d_Y, t, static_cast<hipStream_t>(stream));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_spmm_typed(spops_matrix A, int64_t k, double alpha, const void *d_X, int x_type, double beta,
                                void *d_Y, int y_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    SPOPS_REQUIRE((d_X != nullptr || A->n_cols == 0) && (d_Y != nullptr || A->n_rows == 0), SPOPS_ERROR_INVALID_VALUE,
                  "null X or Y");
    launch_spmm(*A, k, alpha, d_X, x_type, beta, d_Y, y_type, static_cast<hipStream_t>(stream));
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

extern "C" int spops_csr_set_values(spops_matrix A, const void *d_values, int values_type, void *stream) {
  return guarded([&]() {
    require_csr(A);
    require_type(values_type, "values");
    SPOPS_REQUIRE(A->nnz == 0 || d_values != nullptr, SPOPS_ERROR_INVALID_VALUE, "null values");
    const bool perm = A->csr.tperm.p != nullptr;
    convert_values(d_values, values_type, perm ? A->csr.tperm.p : nullptr, A->csr.tperm64, A->nnz, A->csr.val.p,
                   A->value_type, static_cast<hipStream_t>(stream));
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
    std::vector<int2> tasks;
    std::vector<int> longs;
    build_csr_plan(trp, m, tasks, longs);
    P.tasks.upload(tasks);
    P.longs.upload(longs);
    P.ntasks = (int)tasks.size();
    P.nlong = (int)longs.size();
    T->slots = nnz;
    T->dev_bytes = P.rowptr.bytes + P.col.bytes + P.val.bytes + P.tasks.bytes + P.longs.bytes + P.tperm.bytes;
    T->algo_bytes = spmv_algo_bytes(*T, A->index_type == SPOPS_INDEX_I32 ? 4 : 8);
    *AT = T.release();
    return SPOPS_SUCCESS;
  });
}
