// Matrix features of the reference's feature extractor (MTX<I,V>, include/sparse_features.h:
// 13-326; MtxLoad src/sparse_features.cpp:170-480, CalculateFeatures :643-760,
// CalculateTilesFeatures :763-850, P_ratioAndGini :609-640), for a general (not symmetric-
// stored) real CSR matrix.  They are the inputs of the batched SpMM autotuner.
//
// The O(nnz) counting runs on the device in one pass, one wave per row: column counts, the
// bitmap of occupied diagonals, entries near the diagonal, lower / upper / diagonal counts,
// the 256 x 256 tile histogram (uneven split: the first n % 256 tiles one row / column
// larger, :186-190, :357-358), distinct rows and distinct columns per tile, diagonal
// dominance per row, the largest |value| off and on the diagonal, and the row / column
// "variability" (max log10|v| - min log10|v|, with the reference's zero-initialised
// extremes, :430-467).  The finalize on the host follows the reference's arithmetic order
// and quirks exactly (tile variance and nonzero ratio start from -1.0, :788-830; P-ratio and
// Gini from the sorted counts with the trapezoidal rule, :89-100, :609-640).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "common.hpp"

namespace spops {

namespace {

constexpr int kTiles = 256;                       // MAT_TILE_SIZE (include/general_config.h:9)
constexpr int kCnt = 8;                           // device counters
enum { C_CLOSE, C_LOWER, C_UPPER, C_DIAG, C_DOM, C_MAXOFF, C_MAXDIAG, C_ROWVAR };

struct TileMap {
  int64_t num, mod, threshold;                    // rows (or columns) per tile, remainder, first of the small tiles
  __host__ __device__ int64_t operator()(int64_t i) const {
    return i < threshold ? i / (num + 1) : mod + (i - threshold) / num;
  }
};

TileMap tile_map(int64_t n) {
  TileMap t;
  t.num = n / kTiles;
  t.mod = n % kTiles;
  t.threshold = t.mod * (t.num + 1);
  return t;
}

// unsigned 64-bit views of non-negative doubles order like the doubles
__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

template <typename I, typename V>
__global__ __launch_bounds__(kBlock) void k_features(const I *__restrict__ rowptr, const I *__restrict__ col,
                                                     const V *__restrict__ val, int64_t n_rows, int64_t n_cols,
                                                     TileMap tr, TileMap tc, int64_t close_threshold,
                                                     unsigned *__restrict__ colcnt, unsigned *__restrict__ diagbits,
                                                     unsigned *__restrict__ tiles, unsigned *__restrict__ uniq_rb,
                                                     unsigned *__restrict__ uniq_cb, unsigned *__restrict__ colbits,
                                                     unsigned long long *__restrict__ colmax,
                                                     unsigned long long *__restrict__ colmin,
                                                     unsigned long long *__restrict__ cnt) {
  __shared__ unsigned rowtiles[kBlock / kWave][kTiles / 32];
  const int lane = lane_id(), w = wave_id();
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t r = (int64_t)blockIdx.x * (kBlock / kWave) + w; r < n_rows; r += stride) {
    if (lane < kTiles / 32) rowtiles[w][lane] = 0u;
    const int64_t p0 = (int64_t)rowptr[r], p1 = (int64_t)rowptr[r + 1];
    const int64_t trow = tr(r);
    unsigned long long close = 0, lower = 0, upper = 0, diag = 0;
    double dom = 0.0, maxoff = -1.0, maxdiag = -1.0, vmax = 0.0, vminpos = INFINITY;
    __builtin_amdgcn_wave_barrier();
    for (int64_t p = p0 + lane; p < p1; p += kWave) {
      const int64_t c = (int64_t)col[p];
      const double a = fabs((double)val[p]);
      const int64_t off = c - r;
      atomicAdd(&colcnt[c], 1u);
      const int64_t d = off + n_rows - 1;
      atomicOr(&diagbits[d >> 5], 1u << (d & 31));
      close += (off <= close_threshold && -off <= close_threshold) ? 1 : 0;
      if (off == 0) {
        ++diag;
        dom += a;
        maxdiag = fmax(maxdiag, a);
      } else {
        (off < 0 ? lower : upper) += 1;
        dom -= a;
        maxoff = fmax(maxoff, a);
      }
      vmax = fmax(vmax, a);
      if (a > 0.0) vminpos = fmin(vminpos, a);
      const int64_t tcol = tc(c);
      atomicAdd(&tiles[trow * kTiles + tcol], 1u);
      atomicOr(&rowtiles[w][tcol >> 5], 1u << (tcol & 31));
      const unsigned bit = 1u << (trow & 31);
      if (!(atomicOr(&colbits[c * (kTiles / 32) + (trow >> 5)], bit) & bit)) atomicAdd(&uniq_cb[trow * kTiles + tcol], 1u);
      if (a > 0.0) {                       // column extremes of |v| (for the column variability)
        atomicMax(&colmax[c], dbits(a));
        atomicMin(&colmin[c], dbits(a));
      }
    }
    __builtin_amdgcn_wave_barrier();
    // distinct rows per tile: every tile column this row touches
    if (lane < kTiles / 32) {
      unsigned m = rowtiles[w][lane];
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        atomicAdd(&uniq_rb[trow * kTiles + lane * 32 + b], 1u);
      }
    }
    // wave reductions
    for (int s = kWave / 2; s > 0; s >>= 1) {
      close += __shfl_xor(close, s, kWave);
      lower += __shfl_xor(lower, s, kWave);
      upper += __shfl_xor(upper, s, kWave);
      diag += __shfl_xor(diag, s, kWave);
      dom += __shfl_xor(dom, s, kWave);
      maxoff = fmax(maxoff, __shfl_xor(maxoff, s, kWave));
      maxdiag = fmax(maxdiag, __shfl_xor(maxdiag, s, kWave));
      vmax = fmax(vmax, __shfl_xor(vmax, s, kWave));
      vminpos = fmin(vminpos, __shfl_xor(vminpos, s, kWave));
    }
    if (lane == 0) {
      if (close) atomicAdd(&cnt[C_CLOSE], close);
      if (lower) atomicAdd(&cnt[C_LOWER], lower);
      if (upper) atomicAdd(&cnt[C_UPPER], upper);
      if (diag) atomicAdd(&cnt[C_DIAG], diag);
      if (dom > 0.0) atomicAdd(&cnt[C_DOM], 1ull);
      if (maxoff >= 0.0) atomicMax(&cnt[C_MAXOFF], dbits(maxoff) + 1ull);     // +1: 0 means none
      if (maxdiag >= 0.0) atomicMax(&cnt[C_MAXDIAG], dbits(maxdiag) + 1ull);
      // row variability: max(0, log10 max|v|) - min(0, log10 min|v|), rows whose minimum is below 1
      if (vminpos < 1.0) {
        const double t = (vmax > 0.0 ? fmax(0.0, log10(vmax)) : 0.0) - log10(vminpos);
        atomicMax(&cnt[C_ROWVAR], dbits(t) + 1ull);
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_col_variability(const unsigned long long *__restrict__ colmax,
                                                            const unsigned long long *__restrict__ colmin,
                                                            int64_t n_cols, unsigned long long *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  unsigned long long best = 0;
  for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < n_cols; c += stride) {
    const double mn = __longlong_as_double((long long)colmin[c]);
    if (colmax[c] == 0ull || !(mn < 1.0)) continue;
    const double mx = __longlong_as_double((long long)colmax[c]);
    const double t = fmax(0.0, log10(mx)) - log10(mn);
    best = max(best, dbits(t) + 1ull);
  }
  for (int s = kWave / 2; s > 0; s >>= 1) best = max(best, (unsigned long long)__shfl_xor(best, s, kWave));
  if (lane_id() == 0 && best) atomicMax(out, best);
}

__global__ __launch_bounds__(kBlock) void k_fill_u64(unsigned long long *__restrict__ p, int64_t n,
                                                     unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

// P_ratioAndGini (src/sparse_features.cpp:609-640) with the reference's types: counts and
// partial sums as integers, the rest in double.
void p_ratio_gini(std::vector<int64_t> v, int64_t nnz, double &p_ratio, double &gini) {
  std::sort(v.begin(), v.end());
  const int64_t length = (int64_t)v.size();
  int64_t quit = length, step = 0, count = 0;
  double p = 0.0, one_minus_p = 0.0;
  do {
    --quit;
    ++count;
    step += v[quit];
    p = (double)count / length;
    one_minus_p = (double)step / nnz;
  } while (p + one_minus_p < 1.0);
  p_ratio = p;
  std::vector<int64_t> cum(v.size());
  std::partial_sum(v.begin(), v.end(), cum.begin());
  // trapezoidalRule (:89-100): x step = length / size = 1
  const double x_step = (double)length / (double)cum.size();
  double area = (cum[0] + 0) / 2.0;
  for (size_t i = 0; i + 1 < cum.size(); ++i) area += (double)(cum[i] + cum[i + 1]) / 2.0 * x_step;
  const double total = (double)nnz * length / 2.0;
  gini = (total - area) / total;
}

// count statistics of one dimension (rows or columns), CalculateFeatures :664-715
void count_stats(const std::vector<int64_t> &n, int64_t nnz, double *o) {
  const int64_t len = (int64_t)n.size();
  int64_t mn = 100000000, mx = 0, nz = 0;
  const double ave = (double)nnz / len;
  double var = 0.0;
  for (int64_t i = 0; i < len; ++i) {
    mn = std::min(mn, n[i]);
    mx = std::max(mx, n[i]);
    if (n[i]) ++nz;
    const double diff = n[i] - ave;
    var += diff * diff;
  }
  var /= len;
  o[0] = (double)nz / len;
  o[1] = (double)mn;
  o[2] = (double)mx;
  o[3] = ave;
  o[4] = var;
  o[5] = std::sqrt(var);
  p_ratio_gini(n, nnz, o[6], o[7]);
}

// tile statistics, CalculateTilesFeatures :763-830 (the -1.0 starts are the reference's)
void tile_stats(const std::vector<int64_t> &n, int64_t nnz, double ave, double *o) {
  double nzr = -1.0, var = -1.0;
  int64_t mn = 100000000, mx = 0;
  for (size_t i = 0; i < n.size(); ++i) {
    if (n[i]) ++nzr;
    const double diff = n[i] - ave;
    var += diff * diff;
    mn = std::min(mn, n[i]);
    mx = std::max(mx, n[i]);
  }
  nzr /= (double)n.size();
  var /= (double)n.size();
  o[0] = nzr;
  o[1] = (double)mn;
  o[2] = (double)mx;
  o[3] = ave;
  o[4] = var;
  o[5] = std::sqrt(std::max(var, 0.0));
  p_ratio_gini(n, nnz, o[6], o[7]);
}

double from_bits(unsigned long long b, double none) {
  if (b == 0) return none;
  double d;
  const unsigned long long x = b - 1;
  std::memcpy(&d, &x, 8);
  return d;
}

template <typename I, typename V>
void features_device(const void *d_rowptr, const void *d_col, const void *d_val, int64_t n_rows, int64_t n_cols,
                     double *out, hipStream_t s) {
  const TileMap tr = tile_map(n_rows), tc = tile_map(n_cols);
  const int64_t ndiag = n_rows + n_cols - 1;
  const int64_t close_threshold = ndiag / 20;
  DevBuf colcnt, diagbits, tiles, urb, ucb, colbits, colmax, colmin, cnt;
  colcnt.alloc((size_t)n_cols * 4);
  diagbits.alloc((size_t)((ndiag + 31) / 32) * 4);
  tiles.alloc((size_t)kTiles * kTiles * 4);
  urb.alloc((size_t)kTiles * kTiles * 4);
  ucb.alloc((size_t)kTiles * kTiles * 4);
  colbits.alloc((size_t)n_cols * (kTiles / 32) * 4);
  colmax.alloc((size_t)n_cols * 8);
  colmin.alloc((size_t)n_cols * 8);
  cnt.alloc(16 * 8);
  for (DevBuf *b : {&colcnt, &diagbits, &tiles, &urb, &ucb, &colbits, &colmax, &cnt})
    SPOPS_HIP_CHECK(hipMemsetAsync(b->p, 0, b->bytes, s));
  hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(kBlock), 0, s, colmin.as<unsigned long long>(), n_cols,
                     (unsigned long long)0x7FF0000000000000ull);     // +inf
  const int64_t waves = std::max<int64_t>(n_rows, 1);
  const unsigned grid = (unsigned)std::min<int64_t>((waves + 3) / 4, 65536);
  if (n_rows > 0)
    hipLaunchKernelGGL((k_features<I, V>), dim3(grid), dim3(kBlock), 0, s, static_cast<const I *>(d_rowptr),
                       static_cast<const I *>(d_col), static_cast<const V *>(d_val), n_rows, n_cols, tr, tc,
                       close_threshold, colcnt.as<unsigned>(), diagbits.as<unsigned>(), tiles.as<unsigned>(),
                       urb.as<unsigned>(), ucb.as<unsigned>(), colbits.as<unsigned>(),
                       colmax.as<unsigned long long>(), colmin.as<unsigned long long>(), cnt.as<unsigned long long>());
  SPOPS_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_col_variability, dim3(1024), dim3(kBlock), 0, s, colmax.as<unsigned long long>(),
                     colmin.as<unsigned long long>(), n_cols, cnt.as<unsigned long long>() + kCnt);
  SPOPS_HIP_CHECK(hipGetLastError());
  // host side
  std::vector<I> hrp(n_rows + 1);
  std::vector<unsigned> hcc(n_cols), hdb((size_t)((ndiag + 31) / 32)), ht(kTiles * kTiles), hu(kTiles * kTiles),
      hv(kTiles * kTiles);
  unsigned long long hc[16];
  SPOPS_HIP_CHECK(hipMemcpyAsync(hrp.data(), d_rowptr, (size_t)(n_rows + 1) * sizeof(I), hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hcc.data(), colcnt.p, (size_t)n_cols * 4, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hdb.data(), diagbits.p, hdb.size() * 4, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(ht.data(), tiles.p, ht.size() * 4, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hu.data(), urb.p, hu.size() * 4, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hv.data(), ucb.p, hv.size() * 4, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hc, cnt.p, sizeof(hc), hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  const int64_t nnz = (int64_t)hrp[n_rows] - (int64_t)hrp[0];
  std::vector<int64_t> rowlen(n_rows), collen(hcc.begin(), hcc.end());
  for (int64_t r = 0; r < n_rows; ++r) rowlen[r] = (int64_t)hrp[r + 1] - (int64_t)hrp[r];
  int64_t ndiags = 0;
  for (unsigned wv : hdb) ndiags += __builtin_popcount(wv);
  out[0] = (double)n_rows;
  out[1] = (double)n_cols;
  out[2] = (double)nnz;
  out[3] = (double)nnz / ((double)n_rows * n_cols);
  out[4] = (double)ndiags;
  out[5] = (double)(int64_t)hc[C_CLOSE] / nnz;
  out[6] = (double)hc[C_LOWER];
  out[7] = (double)hc[C_UPPER];
  out[8] = (double)hc[C_DIAG];
  count_stats(rowlen, nnz, out + 9);
  count_stats(collen, nnz, out + 17);
  out[25] = (double)(int64_t)hc[C_DOM] / (double)std::min(n_rows, n_cols);
  out[26] = from_bits(hc[C_MAXOFF], -std::numeric_limits<double>::max());
  out[27] = from_bits(hc[C_MAXDIAG], -std::numeric_limits<double>::max());
  out[28] = from_bits(hc[C_ROWVAR], -1.0);
  out[29] = from_bits(hc[kCnt], -1.0);
  out[30] = (double)kTiles;
  out[31] = (double)tr.num;
  out[32] = (double)tc.num;
  std::vector<int64_t> t(ht.begin(), ht.end()), rb(kTiles, 0), cb(kTiles, 0);
  for (int i = 0; i < kTiles; ++i)
    for (int j = 0; j < kTiles; ++j) {
      rb[i] += t[i * kTiles + j];
      cb[j] += t[i * kTiles + j];
    }
  tile_stats(t, nnz, (double)nnz / ((double)kTiles * kTiles), out + 33);
  tile_stats(rb, nnz, (double)nnz / kTiles, out + 41);
  tile_stats(cb, nnz, (double)nnz / kTiles, out + 49);
  int64_t su = 0, sv = 0;
  for (unsigned x : hu) su += x;
  for (unsigned x : hv) sv += x;
  out[57] = (double)su / nnz;
  out[58] = (double)sv / nnz;
  out[59] = (double)su / n_rows;
  out[60] = (double)sv / n_cols;
}

}  // namespace

void row_features_host(const int64_t *rowptr, const int64_t *col, int64_t n_rows, int64_t n_cols, double out[3]) {
  const int64_t nnz = rowptr[n_rows] - rowptr[0];
  std::vector<int64_t> len(n_rows);
  for (int64_t r = 0; r < n_rows; ++r) len[r] = rowptr[r + 1] - rowptr[r];
  double st[8];
  count_stats(len, nnz, st);
  const int64_t thr = (n_rows + n_cols - 1) / 20;
  int64_t close = 0;
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) close += std::llabs(col[p] - r) <= thr ? 1 : 0;
  out[0] = st[3];                                     // ave_nnz_each_row
  out[1] = st[7];                                     // Gini_coeff_row
  out[2] = nnz ? (double)close / nnz : 1.0;           // nnz_closed2diag_ratio
}

}  // namespace spops

using namespace spops;

extern "C" int spops_matrix_features(const void *d_rowptr, const void *d_col, const void *d_values, int index_type,
                                     int value_type, int64_t n_rows, int64_t n_cols, double *out, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(n_rows > 0 && n_cols > 0 && d_rowptr && d_col && d_values, SPOPS_ERROR_INVALID_VALUE,
                  "features need a non-empty matrix and its device arrays");
    SPOPS_REQUIRE(index_type == SPOPS_INDEX_I32 || index_type == SPOPS_INDEX_I64, SPOPS_ERROR_INVALID_VALUE,
                  "bad index_type");
    SPOPS_REQUIRE(value_type == SPOPS_VALUE_F64 || value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "values must be F64 or F32");
    SPOPS_REQUIRE(n_cols < ((int64_t)1 << 31) && n_rows < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT,
                  "rows and columns must be < 2^31");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (index_type == SPOPS_INDEX_I32) {
      if (value_type == SPOPS_VALUE_F64) features_device<int32_t, double>(d_rowptr, d_col, d_values, n_rows, n_cols, out, s);
      else features_device<int32_t, float>(d_rowptr, d_col, d_values, n_rows, n_cols, out, s);
    } else {
      if (value_type == SPOPS_VALUE_F64) features_device<int64_t, double>(d_rowptr, d_col, d_values, n_rows, n_cols, out, s);
      else features_device<int64_t, float>(d_rowptr, d_col, d_values, n_rows, n_cols, out, s);
    }
    return SPOPS_SUCCESS;
  });
}
