// Device-wide scan and stable LSD radix sort for gfx950 (wave64).
#include "primitives.hpp"

namespace spops {

namespace {

constexpr int kScanItems = 8;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;      // 2048
constexpr int kSortItems = 16;                      // per thread
constexpr int kSortTile = kBlock * kSortItems;      // 4096
constexpr int kRadix = 256;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int64_t o = __shfl_up(v, d, kWave);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of the kBlock values `v` (one per thread); returns the block total too.
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *lds4, int64_t &total) {
  int64_t inc = wave_incl_scan(v);
  if (lane_id() == kWave - 1) lds4[wave_id()] = inc;
  __syncthreads();
  int64_t wofs = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    int64_t s = lds4[w];
    if (w < wave_id()) wofs += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wofs + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_tile_sum(const int64_t *__restrict__ in, int64_t n, int64_t *__restrict__ sums) {
  __shared__ int64_t lds4[kBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
    if (j < n) s += in[j];
  }
  int64_t total;
  block_excl_scan(s, lds4, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_tile_scan(const int64_t *__restrict__ in, int64_t *__restrict__ out, int64_t n,
                                                      const int64_t *__restrict__ offs) {
  __shared__ int64_t tile[kScanTile];
  __shared__ int64_t lds4[kBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
    tile[i * kBlock + threadIdx.x] = j < n ? in[j] : 0;
  }
  __syncthreads();
  int64_t loc[kScanItems], s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    loc[i] = s;
    s += tile[threadIdx.x * kScanItems + i];
  }
  int64_t total;
  int64_t pre = block_excl_scan(s, lds4, total) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) tile[threadIdx.x * kScanItems + i] = pre + loc[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
    if (j < n) out[j] = tile[i * kBlock + threadIdx.x];
  }
}

__device__ __forceinline__ uint32_t digit_of(uint64_t key, int shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
}

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                       uint32_t mask, int64_t ntiles, int64_t *__restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; ++i) {
    int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
    if (j < n) atomicAdd(&h[digit_of(keys[j], shift, mask)], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <typename VAL>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t *__restrict__ kin, const VAL *__restrict__ vin,
                                                          uint64_t *__restrict__ kout, VAL *__restrict__ vout, int64_t n,
                                                          int shift, uint32_t mask, int64_t ntiles,
                                                          const int64_t *__restrict__ hist) {
  __shared__ int64_t gofs[kRadix];
  __shared__ int run[kRadix];
  __shared__ int wcnt[kBlock / kWave][kRadix];
  gofs[threadIdx.x] = hist[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  run[threadIdx.x] = 0;
  const int lane = lane_id(), w = wave_id();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int it = 0; it < kSortItems; ++it) {
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q) wcnt[q][threadIdx.x] = 0;
    const int64_t j = base + (int64_t)it * kBlock + threadIdx.x;
    const bool valid = j < n;
    uint64_t key = 0;
    VAL val{};
    if (valid) {
      key = kin[j];
      if (vin) val = vin[j];
    }
    const uint32_t d = valid ? digit_of(key, shift, mask) : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(valid && bit);
      peers &= bit ? bal : ~bal;
    }
    const int lower = __popcll(peers & lt_mask);
    const int cnt = __popcll(peers);
    __syncthreads();
    if (valid && lower == 0) wcnt[w][d] = cnt;
    __syncthreads();
    if (valid) {
      int pre = run[d] + lower;
      for (int q = 0; q < w; ++q) pre += wcnt[q][d];
      const int64_t dst = gofs[d] + pre;
      kout[dst] = key;
      if (vout) vout[dst] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q) add += wcnt[q][threadIdx.x];
    run[threadIdx.x] += add;
    __syncthreads();
  }
}

struct Tmp {
  DevBuf b;
  void *get(size_t n) {
    if (b.bytes < n) b.alloc(n);
    return b.p;
  }
};

}  // namespace

void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s, int64_t *total) {
  if (n <= 0) {
    if (total) *total = 0;
    return;
  }
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  DevBuf sums;
  sums.alloc((size_t)(ntiles + 1) * sizeof(int64_t));
  int64_t *ps = sums.as<int64_t>();
  int64_t last_in = 0;
  if (total) SPOPS_HIP_CHECK(hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, s));
  if (total) SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  if (ntiles > 1) {
    hipLaunchKernelGGL(k_tile_sum, dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, n, ps);
    SPOPS_HIP_CHECK(hipGetLastError());
    exclusive_scan_i64(ps, ps, ntiles, s, nullptr);
  }
  hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, out, n, ntiles > 1 ? ps : nullptr);
  SPOPS_HIP_CHECK(hipGetLastError());
  if (total) {
    int64_t last_out = 0;
    SPOPS_HIP_CHECK(hipMemcpyAsync(&last_out, out + n - 1, 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    *total = last_out + last_in;
  }
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));   // `sums` is freed on return
}

void radix_sort_u64(uint64_t *keys, void *vals, int val_bytes, int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1 || end_bit <= begin_bit) return;
  SPOPS_REQUIRE(vals == nullptr || val_bytes == 4 || val_bytes == 8, SPOPS_ERROR_INVALID_VALUE, "payload must be 4 or 8 bytes");
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  DevBuf tk, tv, hist;
  tk.alloc((size_t)n * 8);
  if (vals) tv.alloc((size_t)n * val_bytes);
  hist.alloc((size_t)kRadix * ntiles * 8);
  uint64_t *kin = keys, *kout = tk.as<uint64_t>();
  void *vin = vals, *vout = vals ? tv.p : nullptr;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
    const int bits = std::min(8, end_bit - shift);
    const uint32_t mask = (1u << bits) - 1u;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kBlock), 0, s, kin, n, shift, mask, ntiles,
                       hist.as<int64_t>());
    SPOPS_HIP_CHECK(hipGetLastError());
    exclusive_scan_i64(hist.as<int64_t>(), hist.as<int64_t>(), kRadix * ntiles, s, nullptr);
    if (!vals || val_bytes == 4)
      hipLaunchKernelGGL(k_radix_scatter<uint32_t>, dim3((unsigned)ntiles), dim3(kBlock), 0, s, kin,
                         static_cast<const uint32_t *>(vin), kout, static_cast<uint32_t *>(vout), n, shift, mask, ntiles,
                         hist.as<int64_t>());
    else
      hipLaunchKernelGGL(k_radix_scatter<uint64_t>, dim3((unsigned)ntiles), dim3(kBlock), 0, s, kin,
                         static_cast<const uint64_t *>(vin), kout, static_cast<uint64_t *>(vout), n, shift, mask, ntiles,
                         hist.as<int64_t>());
    SPOPS_HIP_CHECK(hipGetLastError());
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (kin != keys) {
    SPOPS_HIP_CHECK(hipMemcpyAsync(keys, kin, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    if (vals) SPOPS_HIP_CHECK(hipMemcpyAsync(vals, vin, (size_t)n * val_bytes, hipMemcpyDeviceToDevice, s));
  }
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));   // temporaries are freed on return
}

}  // namespace spops
