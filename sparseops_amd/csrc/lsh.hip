// MinHash signatures and LSH candidate pairs on the GPU, bit-exact with the reference:
//   minhash_signatures_flat            include/spgemm_MinHashLSH.h:119-180
//   band_hash                          :277-283
//   lsh_candidate_pairs_from_flat_vector (records path :444-543, bucket limit 256,
//                                      window 16, :235-243)
//   minhash_estimated_jaccard          :183-204
//
// Pipeline (all on the device):
//   1. k_minhash: one wave per row; lane d keeps min_j (a_d*col_j + b_d) mod 2^64 for its
//      dimension(s); columns are broadcast from a 64-column register tile with readlane.
//   2. k_band_records: key = band hash of (row, band), payload = band<<32 | row.
//   3. stable radix sort by the 64-bit bucket, then a stable pass on the band bits of the
//      payload -> records ordered by (band, bucket, row), rows ascending inside a bucket.
//   4. bucket heads -> bucket ids (scan) -> per-record pair counts (all later members, or
//      the next <= 16 when the bucket has > 256 members and constrain is set) -> scan ->
//      k_emit writes (i << 32 | j) with i < j.
//   5. radix sort of the pair keys, unique, and k_score: one wave per pair, ballot of equal
//      signature dimensions, score = matches / k.
#include <memory>

#include "lsh_internal.hpp"
#include "primitives.hpp"

namespace spops {

namespace {

template <typename I>
__global__ __launch_bounds__(kBlock) void k_minhash(const I *__restrict__ rowptr, const I *__restrict__ col, int64_t n,
                                                    int k, const uint64_t *__restrict__ coef, uint64_t *__restrict__ sigs) {
  const int64_t row = __builtin_amdgcn_readfirstlane((int)blockIdx.x) * (int64_t)(kBlock / kWave) + wave_id();
  if (row >= n) return;
  const int lane = lane_id();
  uint64_t a[kMaxMinhashDims / kWave], b[kMaxMinhashDims / kWave], m[kMaxMinhashDims / kWave];
#pragma unroll
  for (int t = 0; t < kMaxMinhashDims / kWave; ++t) {
    const int d = lane + t * kWave;
    a[t] = d < k ? coef[d] : 0;
    b[t] = d < k ? coef[k + d] : 0;
    m[t] = ~0ull;
  }
  const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
  const int nt = (k + kWave - 1) / kWave;
  for (int64_t base = p0; base < p1; base += kWave) {
    const uint64_t cl = (base + lane < p1) ? (uint64_t)col[base + lane] : 0ull;
    const uint32_t lo = (uint32_t)cl, hi = (uint32_t)(cl >> 32);
    const int cnt = (int)((p1 - base) < kWave ? (p1 - base) : kWave);
    for (int j = 0; j < cnt; ++j) {
      const uint64_t xj = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)lo, j) |
                          ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, j) << 32);
#pragma unroll
      for (int t = 0; t < kMaxMinhashDims / kWave; ++t) {
        if (t < nt) {
          const uint64_t h = a[t] * xj + b[t];
          m[t] = h < m[t] ? h : m[t];
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kMaxMinhashDims / kWave; ++t) {
    const int d = lane + t * kWave;
    if (d < k) sigs[row * k + d] = m[t];
  }
}

__global__ __launch_bounds__(kBlock) void k_band_records(const uint64_t *__restrict__ sigs, int64_t n, int k, int r,
                                                         int bands, uint64_t *__restrict__ keys,
                                                         uint64_t *__restrict__ payload) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n * bands) return;
  const int band = (int)(idx / n);
  const int64_t row = idx - (int64_t)band * n;
  const uint64_t *s = sigs + row * k + (int64_t)band * r;
  uint64_t h = 0;
  for (int i = 0; i < r; ++i) h = h * 0x9e3779b97f4a7c15ull + s[i];
  keys[idx] = h;
  payload[idx] = ((uint64_t)band << 32) | (uint64_t)row;
}

// after the band pass the arrays are swapped: key = band<<32|row, payload = bucket
__global__ __launch_bounds__(kBlock) void k_heads(const uint64_t *__restrict__ brow, const uint64_t *__restrict__ bucket,
                                                  int64_t m, int64_t *__restrict__ head) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  head[a] = (a == 0 || (brow[a] >> 32) != (brow[a - 1] >> 32) || bucket[a] != bucket[a - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_group_starts(const int64_t *__restrict__ head_scan, const int64_t *__restrict__ head,
                                                         int64_t m, int64_t *__restrict__ gstart) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  if (head[a]) gstart[head_scan[a]] = a;
}

__global__ __launch_bounds__(kBlock) void k_pair_counts(const int64_t *__restrict__ head_scan, const int64_t *__restrict__ head,
                                                        const int64_t *__restrict__ gstart, int64_t ngroups, int64_t m,
                                                        int constrain, int64_t *__restrict__ cnt,
                                                        int64_t *__restrict__ gid_out) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  const int64_t g = head_scan[a] + head[a] - 1;     // bucket id of record a
  const int64_t e = (g + 1 < ngroups) ? gstart[g + 1] : m;
  const int64_t size = e - gstart[g];
  int64_t c = e - a - 1;
  if (constrain && size > kLshBucketLimit && c > kLshWindow) c = kLshWindow;
  cnt[a] = c;
  gid_out[a] = g;
}

__global__ __launch_bounds__(kBlock) void k_emit(const uint64_t *__restrict__ brow, const int64_t *__restrict__ ofs,
                                                 const int64_t *__restrict__ cnt, int64_t m, uint64_t *__restrict__ pairs) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  const uint64_t i = brow[a] & 0xffffffffull;
  const int64_t o = ofs[a], c = cnt[a];
  for (int64_t t = 0; t < c; ++t) pairs[o + t] = (i << 32) | (brow[a + 1 + t] & 0xffffffffull);
}

__global__ __launch_bounds__(kBlock) void k_unique_flags(const uint64_t *__restrict__ keys, int64_t m, int64_t *__restrict__ f) {
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (a >= m) return;
  f[a] = (a == 0 || keys[a] != keys[a - 1]) ? 1 : 0;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_unique_score(const uint64_t *__restrict__ keys, const int64_t *__restrict__ f,
                                                         const int64_t *__restrict__ pos, int64_t m,
                                                         const uint64_t *__restrict__ sigs, int k,
                                                         int64_t *__restrict__ pi, int64_t *__restrict__ pj,
                                                         S *__restrict__ score) {
  // one wave per raw pair, the grid striding over them (there can be more than 2^26 pairs)
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t a = (int64_t)blockIdx.x * (kBlock / kWave) + wave_id(); a < m; a += stride) {
    if (!f[a]) continue;
    const uint64_t key = keys[a];
    const int64_t i = (int64_t)(key >> 32), j = (int64_t)(key & 0xffffffffull);
    int match = 0;
    for (int d0 = 0; d0 < k; d0 += kWave) {
      const int d = d0 + lane_id();
      const bool eq = d < k && sigs[i * k + d] == sigs[j * k + d];
      match += __popcll(__ballot(eq));
    }
    if (lane_id() == 0) {
      const int64_t o = pos[a];
      pi[o] = i;
      pj[o] = j;
      score[o] = (S)((double)match / (double)k);
    }
  }
}

inline unsigned blocks_for(int64_t n, int per_block) { return grid_blocks(n, per_block); }

}  // namespace

void minhash_coefficients(int k, uint64_t seed, std::vector<uint64_t> &coef) {
  coef.resize(2 * (size_t)k);
  uint64_t s = seed;
  for (int d = 0; d < k; ++d) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    coef[d] = (s >> 1) * 2 + 1;
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    coef[k + d] = s;
  }
}

void minhash_device(const void *rowptr, const void *col, int index_type, int64_t n, int k, uint64_t seed, uint64_t *sigs,
                    hipStream_t s) {
  SPOPS_REQUIRE(k > 0 && k <= kMaxMinhashDims, SPOPS_ERROR_LIMIT, "k must be in [1, %d]", kMaxMinhashDims);
  if (n <= 0) return;
  std::vector<uint64_t> coef;
  minhash_coefficients(k, seed, coef);
  DevBuf dcoef;
  dcoef.upload(coef);
  const unsigned grid = blocks_for(n, kBlock / kWave);
  if (index_type == SPOPS_INDEX_I32)
    hipLaunchKernelGGL(k_minhash<int32_t>, dim3(grid), dim3(kBlock), 0, s, static_cast<const int32_t *>(rowptr),
                       static_cast<const int32_t *>(col), n, k, dcoef.as<uint64_t>(), sigs);
  else
    hipLaunchKernelGGL(k_minhash<int64_t>, dim3(grid), dim3(kBlock), 0, s, static_cast<const int64_t *>(rowptr),
                       static_cast<const int64_t *>(col), n, k, dcoef.as<uint64_t>(), sigs);
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
}

void lsh_pairs_device(const uint64_t *sigs, int64_t n, int k, int bands, int constrain, int score_type,
                      spops_pairs_s &out, hipStream_t s) {
  out.count = 0;
  out.score_type = score_type;
  const int r = bands > 0 ? k / bands : 0;
  if (n <= 0 || bands <= 0 || k <= 0 || r <= 0 || r * bands != k) return;   // reference returns empty
  SPOPS_REQUIRE(n < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "LSH needs n_rows < 2^31");
  const int64_t m = n * bands;
  DevBuf keys, pay;
  keys.alloc((size_t)m * 8);
  pay.alloc((size_t)m * 8);
  hipLaunchKernelGGL(k_band_records, dim3(blocks_for(m, kBlock)), dim3(kBlock), 0, s, sigs, n, k, r, bands,
                     keys.as<uint64_t>(), pay.as<uint64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  radix_sort_u64(keys.as<uint64_t>(), pay.p, 8, m, 0, 64, s);                  // by bucket, stable
  const int band_bits = bit_width_u64((uint64_t)(bands - 1));
  if (band_bits > 0) radix_sort_u64(pay.as<uint64_t>(), keys.p, 8, m, 32, 32 + band_bits, s);  // by band, stable
  // pay = band<<32|row (sorted), keys = bucket
  DevBuf head, hscan, gid, cnt;
  head.alloc((size_t)m * 8);
  hscan.alloc((size_t)m * 8);
  hipLaunchKernelGGL(k_heads, dim3(blocks_for(m, kBlock)), dim3(kBlock), 0, s, pay.as<uint64_t>(), keys.as<uint64_t>(), m,
                     head.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  int64_t ngroups = 0;
  exclusive_scan_i64(head.as<int64_t>(), hscan.as<int64_t>(), m, s, &ngroups);
  DevBuf gstart;
  gstart.alloc((size_t)(ngroups + 1) * 8);
  hipLaunchKernelGGL(k_group_starts, dim3(blocks_for(m, kBlock)), dim3(kBlock), 0, s, hscan.as<int64_t>(),
                     head.as<int64_t>(), m, gstart.as<int64_t>());
  cnt.alloc((size_t)m * 8);
  gid.alloc((size_t)m * 8);
  hipLaunchKernelGGL(k_pair_counts, dim3(blocks_for(m, kBlock)), dim3(kBlock), 0, s, hscan.as<int64_t>(),
                     head.as<int64_t>(), gstart.as<int64_t>(), ngroups, m, constrain, cnt.as<int64_t>(),
                     gid.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  head.release();
  hscan.release();
  gstart.release();
  gid.release();
  DevBuf ofs;
  ofs.alloc((size_t)m * 8);
  int64_t npairs_raw = 0;
  exclusive_scan_i64(cnt.as<int64_t>(), ofs.as<int64_t>(), m, s, &npairs_raw);
  SPOPS_REQUIRE(npairs_raw <= kLshMaxRawPairs, SPOPS_ERROR_LIMIT,
                "LSH would emit %lld candidate pairs before de-duplication (limit %lld); use constrain=1",
                (long long)npairs_raw, (long long)kLshMaxRawPairs);
  DevBuf raw;
  raw.alloc((size_t)std::max<int64_t>(npairs_raw, 1) * 8);
  hipLaunchKernelGGL(k_emit, dim3(blocks_for(m, kBlock)), dim3(kBlock), 0, s, pay.as<uint64_t>(), ofs.as<int64_t>(),
                     cnt.as<int64_t>(), m, raw.as<uint64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  keys.release();
  pay.release();
  cnt.release();
  ofs.release();
  if (npairs_raw == 0) return;
  const int row_bits = bit_width_u64((uint64_t)(n - 1));
  radix_sort_u64(raw.as<uint64_t>(), nullptr, 0, npairs_raw, 0, 32 + row_bits, s);
  DevBuf flag, pos;
  flag.alloc((size_t)npairs_raw * 8);
  pos.alloc((size_t)npairs_raw * 8);
  hipLaunchKernelGGL(k_unique_flags, dim3(blocks_for(npairs_raw, kBlock)), dim3(kBlock), 0, s, raw.as<uint64_t>(),
                     npairs_raw, flag.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  int64_t nu = 0;
  exclusive_scan_i64(flag.as<int64_t>(), pos.as<int64_t>(), npairs_raw, s, &nu);
  out.pi.alloc((size_t)nu * 8);
  out.pj.alloc((size_t)nu * 8);
  out.score.alloc((size_t)nu * (score_type == SPOPS_VALUE_F32 ? 4 : 8));
  const unsigned grid = std::min<unsigned>(blocks_for(std::min<int64_t>(npairs_raw, (int64_t)1 << 24), kBlock / kWave),
                                           1u << 22);
  if (score_type == SPOPS_VALUE_F32)
    hipLaunchKernelGGL(k_unique_score<float>, dim3(grid), dim3(kBlock), 0, s, raw.as<uint64_t>(), flag.as<int64_t>(),
                       pos.as<int64_t>(), npairs_raw, sigs, k, out.pi.as<int64_t>(), out.pj.as<int64_t>(),
                       out.score.as<float>());
  else
    hipLaunchKernelGGL(k_unique_score<double>, dim3(grid), dim3(kBlock), 0, s, raw.as<uint64_t>(), flag.as<int64_t>(),
                       pos.as<int64_t>(), npairs_raw, sigs, k, out.pi.as<int64_t>(), out.pj.as<int64_t>(),
                       out.score.as<double>());
  SPOPS_HIP_CHECK(hipGetLastError());
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  out.count = nu;
  out.raw_count = npairs_raw;
}

}  // namespace spops

using namespace spops;

extern "C" int spops_minhash_signatures(const void *d_rowptr, const void *d_col, int index_type, int64_t n_rows, int k,
                                        uint64_t seed, uint64_t *d_sigs, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(n_rows >= 0, SPOPS_ERROR_INVALID_VALUE, "negative n_rows");
    SPOPS_REQUIRE(n_rows == 0 || (d_rowptr && d_col && d_sigs), SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_REQUIRE(index_type == SPOPS_INDEX_I32 || index_type == SPOPS_INDEX_I64, SPOPS_ERROR_INVALID_VALUE,
                  "bad index_type");
    minhash_device(d_rowptr, d_col, index_type, n_rows, k, seed, d_sigs, static_cast<hipStream_t>(stream));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lsh_pairs_from_signatures(const uint64_t *d_sigs, int64_t n_rows, int k, int bands, int constrain,
                                               int score_type, spops_pairs *out, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(score_type == SPOPS_VALUE_F64 || score_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "score_type must be F64 or F32");
    auto P = std::make_unique<spops_pairs_s>();
    lsh_pairs_device(d_sigs, n_rows, k, bands, constrain, score_type, *P, static_cast<hipStream_t>(stream));
    *out = P.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_lsh_candidate_pairs(const void *d_rowptr, const void *d_col, int index_type, int64_t n_rows, int k,
                                         int bands, uint64_t seed, int constrain, int score_type, spops_pairs *out,
                                         void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(out != nullptr, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(score_type == SPOPS_VALUE_F64 || score_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "score_type must be F64 or F32");
    SPOPS_REQUIRE(k > 0 && k <= kMaxMinhashDims, SPOPS_ERROR_LIMIT, "k must be in [1, %d]", kMaxMinhashDims);
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf sigs;
    sigs.alloc((size_t)std::max<int64_t>(n_rows, 1) * k * 8);
    minhash_device(d_rowptr, d_col, index_type, n_rows, k, seed, sigs.as<uint64_t>(), s);
    auto P = std::make_unique<spops_pairs_s>();
    lsh_pairs_device(sigs.as<uint64_t>(), n_rows, k, bands, constrain, score_type, *P, s);
    *out = P.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_pairs_count(spops_pairs P, int64_t *count) {
  return guarded([&]() {
    SPOPS_REQUIRE(P && count, SPOPS_ERROR_INVALID_VALUE, "null argument");
    *count = P->count;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_pairs_device(spops_pairs P, const int64_t **d_i, const int64_t **d_j, const void **d_score) {
  return guarded([&]() {
    SPOPS_REQUIRE(P, SPOPS_ERROR_INVALID_VALUE, "null pairs");
    if (d_i) *d_i = P->pi.as<int64_t>();
    if (d_j) *d_j = P->pj.as<int64_t>();
    if (d_score) *d_score = P->score.p;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_pairs_download(spops_pairs P, int64_t *i, int64_t *j, void *score) {
  return guarded([&]() {
    SPOPS_REQUIRE(P, SPOPS_ERROR_INVALID_VALUE, "null pairs");
    const size_t n = (size_t)P->count;
    if (n == 0) return SPOPS_SUCCESS;
    if (i) SPOPS_HIP_CHECK(hipMemcpy(i, P->pi.p, n * 8, hipMemcpyDeviceToHost));
    if (j) SPOPS_HIP_CHECK(hipMemcpy(j, P->pj.p, n * 8, hipMemcpyDeviceToHost));
    if (score) SPOPS_HIP_CHECK(hipMemcpy(score, P->score.p, n * (P->score_type == SPOPS_VALUE_F32 ? 4 : 8), hipMemcpyDeviceToHost));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_pairs_destroy(spops_pairs P) {
  delete P;
  return SPOPS_SUCCESS;
}
