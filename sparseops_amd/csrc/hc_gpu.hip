// Approximate row clustering on the GPU: a parallel alternative to the sequential hierarchical
// clustering (include/spgemm_utility.h:743-915) for the LeSpGEMM preprocessing.
//
// The candidate pairs (i, j, score) of MinHash + LSH are the edges of a graph over the rows.
// Clusters start as single rows and merge in rounds of mutual best choices:
//   1. every cluster c picks its best edge among the pairs joining it to another cluster d with
//      size(c) + size(d) <= cluster_size: the highest score, ties broken by the smaller hash of the
//      (min(c, d), max(c, d)) key (a bijective mix, so a strict total order on edges: the best
//      valid edge of the whole graph is picked by both of its ends and every round with a valid
//      edge merges; being pseudo-random, it also keeps the long runs of equal-score pairs that the
//      windowed LSH buckets of identical rows produce from merging one pair per round);
//   2. clusters that picked each other merge (the pairs of one round are disjoint), the larger id
//      into the smaller, so a cluster's id is its smallest row;
//   3. rows take their cluster's new id.
// The rounds stop when none merges (or at max_rounds).  Scores between clusters are those of the
// pairs between their rows (single linkage over the candidate graph); the reference instead
// scores discovered root pairs by an exact Jaccard, which is what makes its process sequential.
// Pairs naming a row outside [0, n_rows) are ignored; scores must be >= 0 (Jaccard indices).
// The output follows the reference's convention: clusters ordered by their id (smallest row),
// rows ascending inside.  Deterministic: every choice is a max / min over a set.
#include <algorithm>
#include <vector>

#include "primitives.hpp"

namespace spops {
namespace {

constexpr unsigned long long kNoEdge = ~0ull;

__device__ __forceinline__ unsigned long long edge_key(int64_t a, int64_t b) {
  return (unsigned long long)min(a, b) << 32 | (unsigned long long)max(a, b);
}

// the splitmix64 finalizer: a bijection on 64-bit words
__device__ __forceinline__ unsigned long long rank_hash(unsigned long long x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

template <typename S>
__device__ __forceinline__ unsigned long long score_bits(S s) {
  return (unsigned long long)__double_as_longlong((double)s);     // scores are >= 0: the bits order as the values
}

__global__ __launch_bounds__(kBlock) void k_cm_init(int64_t n, int32_t *__restrict__ label, int32_t *__restrict__ size) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    label[i] = (int32_t)i;
    size[i] = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_cm_reset(int64_t n, unsigned long long *__restrict__ best_s,
                                                     unsigned long long *__restrict__ best_h,
                                                     unsigned long long *__restrict__ best_e) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    best_s[i] = 0;
    best_h[i] = ~0ull;
    best_e[i] = kNoEdge;
  }
}

// pass 1: the best score of every cluster over its valid edges (a score of 0 is a valid edge:
// best_s stores bits + 1 so that 0 means none)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cm_best_score(int64_t n, int64_t np, const int64_t *__restrict__ pi,
                                                          const int64_t *__restrict__ pj, const S *__restrict__ sc,
                                                          const int32_t *__restrict__ label,
                                                          const int32_t *__restrict__ size, int64_t cs,
                                                          unsigned long long *__restrict__ best_s) {
  for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < np; t += (int64_t)gridDim.x * kBlock) {
    if (pi[t] < 0 || pi[t] >= n || pj[t] < 0 || pj[t] >= n) continue;
    const int32_t a = label[pi[t]], b = label[pj[t]];
    if (a == b || (int64_t)size[a] + size[b] > cs) continue;
    const unsigned long long v = score_bits(sc[t]) + 1;
    atomicMax(&best_s[a], v);
    atomicMax(&best_s[b], v);
  }
}

// pass 2 (mode 0): among the edges of a cluster at its best score, the smallest key hash;
// pass 3 (mode 1): the key of the edge with that hash (the hash is a bijection of the key, so all
// writers of a cluster write the same key)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cm_best_edge(int64_t n, int64_t np, const int64_t *__restrict__ pi,
                                                         const int64_t *__restrict__ pj, const S *__restrict__ sc,
                                                         const int32_t *__restrict__ label,
                                                         const int32_t *__restrict__ size, int64_t cs,
                                                         const unsigned long long *__restrict__ best_s,
                                                         unsigned long long *__restrict__ best_h,
                                                         unsigned long long *__restrict__ best_e, int mode) {
  for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < np; t += (int64_t)gridDim.x * kBlock) {
    if (pi[t] < 0 || pi[t] >= n || pj[t] < 0 || pj[t] >= n) continue;
    const int32_t a = label[pi[t]], b = label[pj[t]];
    if (a == b || (int64_t)size[a] + size[b] > cs) continue;
    const unsigned long long v = score_bits(sc[t]) + 1, e = edge_key(a, b), h = rank_hash(e);
    if (mode == 0) {
      if (v == best_s[a]) atomicMin(&best_h[a], h);
      if (v == best_s[b]) atomicMin(&best_h[b], h);
    } else {
      if (v == best_s[a] && h == best_h[a]) best_e[a] = e;
      if (v == best_s[b] && h == best_h[b]) best_e[b] = e;
    }
  }
}

// pass 3: clusters whose best edges coincide merge (the larger id into the smaller)
__global__ __launch_bounds__(kBlock) void k_cm_merge(int64_t n, const unsigned long long *__restrict__ best_s,
                                                     const unsigned long long *__restrict__ best_e,
                                                     int32_t *__restrict__ parent, int32_t *__restrict__ size,
                                                     unsigned long long *__restrict__ merged) {
  for (int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlock) {
    parent[c] = (int32_t)c;
    const unsigned long long e = best_e[c];
    if (best_s[c] == 0 || e == kNoEdge) continue;
    const int64_t lo = (int64_t)(e >> 32), hi = (int64_t)(e & 0xFFFFFFFFull);
    if (c != hi) continue;                         // the larger end acts
    if (best_e[lo] != e) continue;                 // not mutual
    parent[c] = (int32_t)lo;
    atomicAdd(merged, 1ull);
  }
}

__global__ __launch_bounds__(kBlock) void k_cm_grow(int64_t n, const int32_t *__restrict__ parent,
                                                    int32_t *__restrict__ size) {
  for (int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlock)
    if (parent[c] != (int32_t)c) size[parent[c]] += size[c];    // each absorbing cluster takes exactly one
}

__global__ __launch_bounds__(kBlock) void k_cm_relabel(int64_t n, const int32_t *__restrict__ parent,
                                                       int32_t *__restrict__ label, uint64_t *__restrict__ keys) {
  for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBlock) {
    const int32_t l = parent[label[r]];
    label[r] = l;
    keys[r] = (uint64_t)(uint32_t)l << 32 | (uint64_t)r;
  }
}

__global__ __launch_bounds__(kBlock) void k_cm_emit(int64_t n, const uint64_t *__restrict__ keys, int64_t *__restrict__ perm,
                                                    int64_t *__restrict__ head) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    perm[i] = (int64_t)(keys[i] & 0xFFFFFFFFull);
    head[i] = (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32)) ? 1 : 0;
  }
}

unsigned grid_for(int64_t items) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + kBlock - 1) / kBlock, 65536));
}

template <typename S>
void cluster_matching(int64_t n, int64_t np, const int64_t *pi, const int64_t *pj, const S *sc, int64_t cs,
                      int64_t max_rounds, int64_t *h_perm, int64_t *h_offsets, int64_t &nclusters, int64_t &rounds,
                      hipStream_t s) {
  DevBuf label, size, parent, best_s, best_h, best_e, merged, keys, perm, head;
  label.alloc((size_t)n * 4);
  size.alloc((size_t)n * 4);
  parent.alloc((size_t)n * 4);
  best_s.alloc((size_t)n * 8);
  best_h.alloc((size_t)n * 8);
  best_e.alloc((size_t)n * 8);
  merged.alloc(8);
  keys.alloc((size_t)n * 8);
  hipLaunchKernelGGL(k_cm_init, dim3(grid_for(n)), dim3(kBlock), 0, s, n, label.as<int32_t>(), size.as<int32_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  rounds = 0;
  for (; rounds < max_rounds; ++rounds) {
    hipLaunchKernelGGL(k_cm_reset, dim3(grid_for(n)), dim3(kBlock), 0, s, n, best_s.as<unsigned long long>(),
                       best_h.as<unsigned long long>(), best_e.as<unsigned long long>());
    SPOPS_HIP_CHECK(hipMemsetAsync(merged.p, 0, 8, s));
    if (np > 0) {
      hipLaunchKernelGGL(k_cm_best_score<S>, dim3(grid_for(np)), dim3(kBlock), 0, s, n, np, pi, pj, sc,
                         label.as<int32_t>(), size.as<int32_t>(), cs, best_s.as<unsigned long long>());
      for (int mode = 0; mode < 2; ++mode)
        hipLaunchKernelGGL(k_cm_best_edge<S>, dim3(grid_for(np)), dim3(kBlock), 0, s, n, np, pi, pj, sc,
                           label.as<int32_t>(), size.as<int32_t>(), cs, best_s.as<unsigned long long>(),
                           best_h.as<unsigned long long>(), best_e.as<unsigned long long>(), mode);
    }
    hipLaunchKernelGGL(k_cm_merge, dim3(grid_for(n)), dim3(kBlock), 0, s, n, best_s.as<unsigned long long>(),
                       best_e.as<unsigned long long>(), parent.as<int32_t>(), size.as<int32_t>(),
                       merged.as<unsigned long long>());
    hipLaunchKernelGGL(k_cm_grow, dim3(grid_for(n)), dim3(kBlock), 0, s, n, parent.as<int32_t>(), size.as<int32_t>());
    hipLaunchKernelGGL(k_cm_relabel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, parent.as<int32_t>(),
                       label.as<int32_t>(), keys.as<uint64_t>());
    SPOPS_HIP_CHECK(hipGetLastError());
    unsigned long long m = 0;
    SPOPS_HIP_CHECK(hipMemcpyAsync(&m, merged.p, 8, hipMemcpyDeviceToHost, s));
    SPOPS_HIP_CHECK(hipStreamSynchronize(s));
    if (m == 0) break;
  }
  // (every round ends with a relabel, so keys hold the final cluster ids)
  // rows ordered by (cluster id, row); a head where the cluster id changes
  radix_sort_u64(keys.as<uint64_t>(), nullptr, 8, n, 0, 32 + std::max(1, bit_width_u64((uint64_t)n)), s);
  perm.alloc((size_t)n * 8);
  head.alloc((size_t)n * 8);
  hipLaunchKernelGGL(k_cm_emit, dim3(grid_for(n)), dim3(kBlock), 0, s, n, keys.as<uint64_t>(), perm.as<int64_t>(),
                     head.as<int64_t>());
  SPOPS_HIP_CHECK(hipGetLastError());
  std::vector<int64_t> hh((size_t)n);
  SPOPS_HIP_CHECK(hipMemcpyAsync(h_perm, perm.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipMemcpyAsync(hh.data(), head.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  SPOPS_HIP_CHECK(hipStreamSynchronize(s));
  nclusters = 0;
  for (int64_t i = 0; i < n; ++i)
    if (hh[(size_t)i]) h_offsets[nclusters++] = i;
  h_offsets[nclusters] = n;
}

}  // namespace
}  // namespace spops

using namespace spops;

extern "C" int spops_cluster_matching(int64_t n_rows, int64_t npairs, const int64_t *d_pi, const int64_t *d_pj,
                                      const void *d_score, int score_type, int64_t cluster_size, int64_t max_rounds,
                                      int64_t *perm, int64_t *offsets, int64_t *nclusters, int64_t *rounds,
                                      void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(perm && offsets && nclusters && rounds, SPOPS_ERROR_INVALID_VALUE, "null output");
    SPOPS_REQUIRE(n_rows >= 0 && npairs >= 0 && cluster_size >= 1 && max_rounds >= 1, SPOPS_ERROR_INVALID_VALUE,
                  "negative size, cluster_size < 1 or max_rounds < 1");
    SPOPS_REQUIRE(n_rows < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "n_rows must be < 2^31");
    SPOPS_REQUIRE(npairs == 0 || (d_pi && d_pj && d_score), SPOPS_ERROR_INVALID_VALUE, "null pair array");
    SPOPS_REQUIRE(score_type == SPOPS_VALUE_F64 || score_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "score_type must be F64 or F32");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n_rows == 0) {
      offsets[0] = 0;
      *nclusters = 0;
      *rounds = 0;
      return SPOPS_SUCCESS;
    }
    if (score_type == SPOPS_VALUE_F64)
      cluster_matching<double>(n_rows, npairs, d_pi, d_pj, static_cast<const double *>(d_score), cluster_size,
                               max_rounds, perm, offsets, *nclusters, *rounds, s);
    else
      cluster_matching<float>(n_rows, npairs, d_pi, d_pj, static_cast<const float *>(d_score), cluster_size,
                              max_rounds, perm, offsets, *nclusters, *rounds, s);
    return SPOPS_SUCCESS;
  });
}
