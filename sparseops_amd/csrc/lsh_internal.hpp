// LSH candidate pairs object and shared constants.
#pragma once

#include <vector>

#include "common.hpp"

namespace spops {

constexpr int kMaxMinhashDims = 256;           // k <= 256 (4 dimensions per lane)
constexpr int64_t kLshBucketLimit = 256;       // include/spgemm_MinHashLSH.h:241
constexpr int64_t kLshWindow = 16;             // include/spgemm_MinHashLSH.h:243
constexpr int64_t kLshMaxRawPairs = (int64_t)1 << 32;   // pairs before de-duplication

void minhash_coefficients(int k, uint64_t seed, std::vector<uint64_t> &coef);

}  // namespace spops

struct spops_pairs_s {
  spops::DevBuf pi, pj, score;
  int64_t count = 0, raw_count = 0;
  int score_type = SPOPS_VALUE_F64;
};
