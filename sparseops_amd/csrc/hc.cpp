// Hierarchical clustering of candidate pairs into row clusters (host, sequential).
//
// Bit-exact restatement of the reference's merge process:
//   hierarchical_clustering_v0_fast  include/spgemm_utility.h:743-915  (version 0)
//   hierarchical_clustering_v1       include/spgemm_utility.h:609-733  (version 1)
//   hierarchical_clustering_v0       include/spgemm_utility.h:517-606  (the map-based original;
//                                    spops_hierarchical_clustering_v0, with the cluster roots)
// It is inherently sequential (a priority-ordered union-find with on-the-fly exact Jaccard
// for discovered root pairs) and stays on the host; the GPU pipeline around it produces
// the pairs and consumes perm/offsets.
//
// v0_fast order of processing: the initial pairs in descending score, and inside one
// score in descending input position (the reference keeps them in
// map<score, vector, greater> and pops from the back), merged with a max-heap of
// discovered pairs under the comparator (score, (i, j)).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>

#include "common.hpp"

namespace spops {
namespace {

// Open-addressing set of (i, j) pairs.
class PairSet {
 public:
  explicit PairSet(size_t expect) { rehash(std::max<size_t>(1024, 2 * expect)); }
  bool insert(int64_t i, int64_t j) {
    if (2 * (n_ + 1) > cap_) rehash(2 * cap_);
    size_t h = slot(i, j);
    for (;;) {
      if (k_[2 * h] < 0) { k_[2 * h] = i; k_[2 * h + 1] = j; ++n_; return true; }
      if (k_[2 * h] == i && k_[2 * h + 1] == j) return false;
      h = (h + 1) & (cap_ - 1);
    }
  }

 private:
  static uint64_t mix(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
  }
  size_t slot(int64_t i, int64_t j) const { return mix((uint64_t)i * 0x9e3779b97f4a7c15ull ^ mix((uint64_t)j)) & (cap_ - 1); }
  void rehash(size_t c) {
    size_t p2 = 1;
    while (p2 < c) p2 <<= 1;
    std::vector<int64_t> old;
    old.swap(k_);
    size_t oc = cap_;
    cap_ = p2;
    n_ = 0;
    k_.assign(2 * cap_, -1);
    for (size_t q = 0; q < oc; ++q)
      if (old[2 * q] >= 0) insert(old[2 * q], old[2 * q + 1]);
  }
  std::vector<int64_t> k_;
  size_t cap_ = 0, n_ = 0;
};

struct Csr64 {
  const int64_t *rp, *col;
};

// include/spgemm_utility.h:402-429 (sorted rows) / :362-387 (general, set semantics for row a)
double jaccard(const Csr64 &A, int64_t a, int64_t b, bool sorted_unique) {
  const int64_t la = A.rp[a + 1] - A.rp[a], lb = A.rp[b + 1] - A.rp[b];
  int64_t c = 0;
  if (sorted_unique) {
    int64_t ia = A.rp[a], ib = A.rp[b];
    while (ia < A.rp[a + 1] && ib < A.rp[b + 1]) {
      if (A.col[ia] == A.col[ib]) { ++c; ++ia; ++ib; }
      else if (A.col[ia] < A.col[ib]) ++ia;
      else ++ib;
    }
  } else {
    std::vector<int64_t> sa(A.col + A.rp[a], A.col + A.rp[a + 1]);
    std::sort(sa.begin(), sa.end());
    for (int64_t j = A.rp[b]; j < A.rp[b + 1]; ++j)
      if (std::binary_search(sa.begin(), sa.end(), A.col[j])) ++c;
  }
  const int64_t u = la + lb - c;
  return u == 0 ? 0.0 : (1.0 * (double)c / (double)u);
}

template <typename V>
struct Item {
  V s;
  int64_t i, j;
};
template <typename V>
bool item_less(const Item<V> &a, const Item<V> &b) {
  if (a.s != b.s) return a.s < b.s;
  if (a.i != b.i) return a.i < b.i;
  return a.j < b.j;
}
template <typename V>
struct ItemLess {
  bool operator()(const Item<V> &a, const Item<V> &b) const { return item_less(a, b); }
};

// Small open-addressing map from a score's bit pattern to a dense group id (grows by
// doubling; the scores of one clustering take few distinct values).
template <typename V>
class ScoreTable {
 public:
  ScoreTable() { keys_.assign(cap_, kEmpty); vals_.assign(cap_, 0); }
  int32_t find_or_add(V s, int32_t next_id) {
    const uint64_t k = bits(s);
    size_t h = slot(k);
    for (;;) {
      if (keys_[h] == k) return vals_[h];
      if (keys_[h] == kEmpty) {
        keys_[h] = k;
        vals_[h] = next_id;
        if (2 * ++n_ > cap_) grow();
        return next_id;
      }
      h = (h + 1) & (cap_ - 1);
    }
  }

 private:
  static constexpr uint64_t kEmpty = ~0ull;             // no score has this bit pattern here (a NaN payload)
  static uint64_t bits(V s) {
    uint64_t b = 0;
    std::memcpy(&b, &s, sizeof(V));
    return b;
  }
  size_t slot(uint64_t k) const { return (size_t)((k * 0x9e3779b97f4a7c15ull) >> 20) & (cap_ - 1); }
  void grow() {
    std::vector<uint64_t> ok;
    std::vector<int32_t> ov;
    ok.swap(keys_);
    ov.swap(vals_);
    cap_ *= 2;
    keys_.assign(cap_, kEmpty);
    vals_.assign(cap_, 0);
    for (size_t q = 0; q < ok.size(); ++q)
      if (ok[q] != kEmpty) {
        size_t h = slot(ok[q]);
        while (keys_[h] != kEmpty) h = (h + 1) & (cap_ - 1);
        keys_[h] = ok[q];
        vals_[h] = ov[q];
      }
  }
  std::vector<uint64_t> keys_;
  std::vector<int32_t> vals_;
  size_t cap_ = 1024, n_ = 0;
};

int64_t find_root(std::vector<int64_t> &cl, int64_t x) {
  int64_t root = x;
  while (cl[root] != root) root = cl[root];
  while (cl[x] != root) { int64_t nx = cl[x]; cl[x] = root; x = nx; }
  return root;
}

// include/spgemm_utility.h:883-914: clusters ordered by root id, rows ascending inside.
int64_t emit(std::vector<int64_t> &cl, int64_t n, int64_t *perm, int64_t *offsets) {
  std::vector<int64_t> root(n), count(n, 0), cid(n, -1);
  for (int64_t i = 0; i < n; ++i) { root[i] = find_root(cl, i); count[root[i]]++; }
  int64_t nc = 0;
  offsets[0] = 0;
  for (int64_t r = 0; r < n; ++r)
    if (count[r] > 0) { cid[r] = nc; offsets[nc + 1] = offsets[nc] + count[r]; ++nc; }
  std::vector<int64_t> cur(offsets, offsets + nc);
  for (int64_t i = 0; i < n; ++i) perm[cur[cid[root[i]]]++] = i;
  return nc;
}

// std::lower_bound over the input pairs with the comparator of spgemm_utility.h:789-796.
// When the input is sorted by (i, j) (the order that search assumes; LSH emits it so) and
// every i is a row, the search runs inside row i's pairs only, found through a row index:
// same answer, a few cache lines instead of ~24 misses over all pairs.
class InitialPairs {
 public:
  InitialPairs(int64_t n, int64_t np, const int64_t *pi, const int64_t *pj) : np_(np), pi_(pi), pj_(pj) {
    bool sorted = true;
    for (int64_t t = 0; t < np && sorted; ++t) {
      if (pi[t] < 0 || pi[t] >= n) sorted = false;
      else if (t > 0 && (pi[t - 1] > pi[t] || (pi[t - 1] == pi[t] && pj[t - 1] > pj[t]))) sorted = false;
    }
    if (!sorted) return;
    first_.assign((size_t)n + 1, 0);
    for (int64_t t = 0; t < np; ++t) ++first_[(size_t)pi[t] + 1];
    for (int64_t r = 0; r < n; ++r) first_[(size_t)r + 1] += first_[(size_t)r];
  }
  bool has(int64_t i, int64_t j) const {
    int64_t first = 0, len = np_;
    if (!first_.empty()) {
      if (i < 0 || i + 1 >= (int64_t)first_.size()) return false;
      first = first_[(size_t)i];
      len = first_[(size_t)i + 1] - first;
    }
    while (len > 0) {
      const int64_t half = len >> 1, mid = first + half;
      const bool less = pi_[mid] != i ? pi_[mid] < i : pj_[mid] < j;
      if (less) { first = mid + 1; len = len - half - 1; }
      else len = half;
    }
    return first < np_ && pi_[first] == i && pj_[first] == j;
  }

 private:
  int64_t np_;
  const int64_t *pi_, *pj_;
  std::vector<int64_t> first_;     // per row: its first pair (empty: unsorted input)
};

// Whether the rows are strictly sorted (spgemm_utility.h:390), asked per row when a Jaccard
// is computed: the sorted-merge and set-semantics counts agree on rows that are sorted and
// duplicate free, so deciding per pair of rows gives the global flag's results.
class RowOrder {
 public:
  RowOrder(const Csr64 &A, int64_t n) : A_(A), state_((size_t)n, 0) {}
  bool sorted(int64_t r) {
    int8_t &st = state_[(size_t)r];
    if (st == 0) {
      st = 1;
      for (int64_t j = A_.rp[r] + 1; j < A_.rp[r + 1]; ++j)
        if (A_.col[j - 1] >= A_.col[j]) { st = 2; break; }
    }
    return st == 1;
  }

 private:
  const Csr64 &A_;
  std::vector<int8_t> state_;      // 0 unknown, 1 strictly ascending, 2 not
};

template <typename V>
int64_t hc_v0_fast(int64_t n, const Csr64 &A, int64_t np, const int64_t *pi, const int64_t *pj, const V *score,
                   int64_t cs, int64_t *perm, int64_t *offsets) {
  std::vector<int64_t> cl(n), sz(n, 1);
  std::vector<char> valid(n, 1);
  for (int64_t i = 0; i < n; ++i) cl[i] = i;
  // Initial stream: group equal scores (counting sort over the distinct values), scores
  // descending, input position descending inside a score.  The distinct scores are few
  // (multiples of 1/k for MinHash estimates): a small open-addressing table keyed by the
  // score's bits (+0.0 for -0.0) gives each pair its group in one pass.
  ScoreTable<V> tab;
  std::vector<int32_t> gid((size_t)np, -1);                 // per pair: group in first-seen order, -1: dropped
  std::vector<V> distinct;
  std::vector<int64_t> gcount;
  for (int64_t t = 0; t < np; ++t) {
    const int64_t a = std::min(pi[t], pj[t]), b = std::max(pi[t], pj[t]);
    if (a == b || a >= n || b >= n) continue;
    const V sc = score[t] == V(0) ? V(0) : score[t];
    const int32_t g = tab.find_or_add(sc, (int32_t)distinct.size());
    if (g == (int32_t)distinct.size()) { distinct.push_back(sc); gcount.push_back(0); }
    gid[(size_t)t] = g;
    ++gcount[(size_t)g];
  }
  // groups in descending score; bstart over that order
  std::vector<int32_t> rank(distinct.size());
  std::iota(rank.begin(), rank.end(), 0);
  std::sort(rank.begin(), rank.end(), [&](int32_t x, int32_t y) { return distinct[(size_t)x] > distinct[(size_t)y]; });
  std::vector<int64_t> fill(distinct.size());                 // per group (first-seen id): next slot
  int64_t total = 0;
  for (int32_t r : rank) { fill[(size_t)r] = total; total += gcount[(size_t)r]; }
  std::vector<int64_t> order((size_t)total);
  for (int64_t t = np - 1; t >= 0; --t)                       // descending position
    if (gid[(size_t)t] >= 0) order[(size_t)fill[(size_t)gid[(size_t)t]]++] = t;
  std::priority_queue<Item<V>, std::vector<Item<V>>, ItemLess<V>> dyn;
  PairSet discovered((size_t)np / 8 + 1024);
  const InitialPairs initial(n, np, pi, pj);
  RowOrder order_of(A, n);
  size_t next = 0;
  auto initial_item = [&](size_t q) {
    int64_t t = order[q];
    return Item<V>{score[t], std::min(pi[t], pj[t]), std::max(pi[t], pj[t])};
  };
  int nclusters = (int)n;
  while ((next < order.size() || !dyn.empty()) && nclusters > 0) {
    Item<V> top;
    if (dyn.empty()) {
      top = initial_item(next++);
    } else if (next >= order.size()) {
      top = dyn.top();
      dyn.pop();
    } else {
      Item<V> init = initial_item(next);
      if (item_less(dyn.top(), init)) { top = init; ++next; }
      else { top = dyn.top(); dyn.pop(); }
    }
    const int64_t i = top.i, j = top.j;
    if (i >= n || j >= n) continue;
    if (cl[i] == i && cl[j] == j) {
      if (!valid[i] || !valid[j]) continue;
      nclusters--;
      if (sz[i] < sz[j]) {
        cl[i] = j; sz[j] += sz[i];
        if (sz[j] >= cs) { valid[j] = 0; nclusters--; }
      } else {
        cl[j] = i; sz[i] += sz[j];
        if (sz[i] >= cs) { valid[i] = 0; nclusters--; }
      }
    } else {
      const int64_t ri = find_root(cl, i), rj = find_root(cl, j);
      if (!valid[ri] || !valid[rj]) continue;
      if (ri != rj) {
        const int64_t a = std::min(ri, rj), b = std::max(ri, rj);
        if (!initial.has(a, b) && discovered.insert(a, b))
          dyn.push(Item<V>{(V)jaccard(A, a, b, order_of.sorted(a) && order_of.sorted(b)), a, b});
      }
    }
  }
  return emit(cl, n, perm, offsets);
}

template <typename V>
int64_t hc_v1(int64_t n, const Csr64 &A, int64_t np, const int64_t *pi, const int64_t *pj, const V *score, int64_t cs,
              int64_t *perm, int64_t *offsets) {
  std::vector<int64_t> cl(n), sz(n, 1);
  std::vector<char> valid(n, 1);
  for (int64_t i = 0; i < n; ++i) cl[i] = i;
  std::priority_queue<Item<V>, std::vector<Item<V>>, ItemLess<V>> sims;
  PairSet seen((size_t)np + 1024);
  for (int64_t t = 0; t < np; ++t) {
    int64_t a = std::min(pi[t], pj[t]), b = std::max(pi[t], pj[t]);
    if (a == b || a >= n || b >= n) continue;
    if (!seen.insert(a, b)) continue;
    sims.push(Item<V>{score[t], a, b});
  }
  int nclusters = (int)n;
  while (!sims.empty() && nclusters > 0) {
    Item<V> top = sims.top();
    sims.pop();
    const int64_t i = top.i, j = top.j;
    if (i >= n || j >= n) continue;
    const int64_t ri = find_root(cl, i), rj = find_root(cl, j);
    if (ri == rj) continue;
    if (!valid[ri] || !valid[rj]) continue;
    if (cl[i] == i && cl[j] == j) {
      nclusters--;
      if (sz[ri] < sz[rj]) {
        cl[ri] = rj; sz[rj] += sz[ri];
        if (sz[rj] >= cs) { valid[rj] = 0; nclusters--; }
      } else {
        cl[rj] = ri; sz[ri] += sz[rj];
        if (sz[ri] >= cs) { valid[ri] = 0; nclusters--; }
      }
    } else {
      const int64_t a = std::min(ri, rj), b = std::max(ri, rj);
      if (seen.insert(a, b)) sims.push(Item<V>{(V)jaccard(A, a, b, false), a, b});
    }
  }
  return emit(cl, n, perm, offsets);
}

// include/spgemm_utility.h:517-606.  The pairs arrive in the order of the caller's map
// (keys as the map holds them, i > j allowed); each enters the queue as (score, (min, max)),
// and a discovered root pair is scored only when its (min, max) key is not a key of the map
// yet.  A pair whose rows are both roots merges them without the root search (a pair (i, i)
// of a root therefore doubles its size, as in the reference); the loop runs while the queue
// is not empty and the cluster count is not zero.
template <typename V>
int64_t hc_v0_map(int64_t n, const Csr64 &A, int64_t np, const int64_t *pi, const int64_t *pj, const V *score, int64_t cs,
                  int64_t *perm, int64_t *offsets, int64_t *roots) {
  std::vector<int64_t> cl(n), sz(n, 1);
  std::vector<char> valid(n, 1);
  for (int64_t i = 0; i < n; ++i) cl[i] = i;
  std::priority_queue<Item<V>, std::vector<Item<V>>, ItemLess<V>> sims;
  PairSet keys((size_t)np + 1024);
  for (int64_t t = 0; t < np; ++t) {
    keys.insert(pi[t], pj[t]);
    sims.push(Item<V>{score[t], std::min(pi[t], pj[t]), std::max(pi[t], pj[t])});
  }
  int64_t nclusters = n;
  while (!sims.empty() && nclusters != 0) {
    const Item<V> top = sims.top();
    sims.pop();
    int64_t i = top.i, j = top.j;
    if (i >= n || j >= n) continue;
    if (cl[i] == i && cl[j] == j) {
      if (!valid[i] || !valid[j]) continue;
      nclusters--;
      if (sz[i] < sz[j]) {
        cl[i] = j; sz[j] += sz[i];
        if (sz[j] >= cs) { valid[j] = 0; nclusters--; }
      } else {
        cl[j] = i; sz[i] += sz[j];
        if (sz[i] >= cs) { valid[i] = 0; nclusters--; }
      }
    } else {
      i = find_root(cl, i);
      j = find_root(cl, j);
      if (!valid[i] || !valid[j]) continue;
      if (i != j) {
        const int64_t a = std::min(i, j), b = std::max(i, j);
        if (keys.insert(a, b)) sims.push(Item<V>{(V)jaccard(A, a, b, false), a, b});
      }
    }
  }
  const int64_t nc = emit(cl, n, perm, offsets);
  for (int64_t c = 0; c < nc; ++c) roots[c] = find_root(cl, perm[offsets[c]]);   // the key of the reference's map
  return nc;
}

}  // namespace
}  // namespace spops

using namespace spops;

extern "C" int spops_hierarchical_clustering_v0(int64_t n, const int64_t *rowptr, const int64_t *col, int64_t np,
                                                const int64_t *pi, const int64_t *pj, const void *score,
                                                int score_type, int64_t cs, int64_t *perm, int64_t *offsets,
                                                int64_t *roots, int64_t *nclusters) {
  return guarded([&]() {
    SPOPS_REQUIRE(n >= 0 && np >= 0, SPOPS_ERROR_INVALID_VALUE, "negative size");
    SPOPS_REQUIRE(n < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "n_rows must be < 2^31");
    SPOPS_REQUIRE(rowptr && perm && offsets && roots && nclusters && (np == 0 || (pi && pj && score)),
                  SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_REQUIRE(score_type == SPOPS_VALUE_F64 || score_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "score_type must be F64 or F32");
    for (int64_t t = 0; t < np; ++t)
      SPOPS_REQUIRE(pi[t] >= 0 && pj[t] >= 0, SPOPS_ERROR_INVALID_VALUE, "negative row in pair %lld", (long long)t);
    Csr64 A{rowptr, col};
    *nclusters = score_type == SPOPS_VALUE_F64
                     ? hc_v0_map<double>(n, A, np, pi, pj, static_cast<const double *>(score), cs, perm, offsets, roots)
                     : hc_v0_map<float>(n, A, np, pi, pj, static_cast<const float *>(score), cs, perm, offsets, roots);
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_hierarchical_clustering(int version, int64_t n, const int64_t *rowptr, const int64_t *col,
                                             int64_t np, const int64_t *pi, const int64_t *pj, const void *score,
                                             int score_type, int64_t cs, int64_t *perm, int64_t *offsets,
                                             int64_t *nclusters) {
  return guarded([&]() {
    SPOPS_REQUIRE(version == 0 || version == 1, SPOPS_ERROR_INVALID_VALUE, "version must be 0 (v0_fast) or 1 (v1)");
    SPOPS_REQUIRE(n >= 0 && np >= 0, SPOPS_ERROR_INVALID_VALUE, "negative size");
    SPOPS_REQUIRE(n < ((int64_t)1 << 31), SPOPS_ERROR_LIMIT, "n_rows must be < 2^31");
    SPOPS_REQUIRE(rowptr && perm && offsets && nclusters && (np == 0 || (pi && pj && score)),
                  SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_REQUIRE(score_type == SPOPS_VALUE_F64 || score_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "score_type must be F64 or F32");
    Csr64 A{rowptr, col};
    if (score_type == SPOPS_VALUE_F64) {
      const double *s = static_cast<const double *>(score);
      *nclusters = version == 0 ? hc_v0_fast<double>(n, A, np, pi, pj, s, cs, perm, offsets)
                                : hc_v1<double>(n, A, np, pi, pj, s, cs, perm, offsets);
    } else {
      const float *s = static_cast<const float *>(score);
      *nclusters = version == 0 ? hc_v0_fast<float>(n, A, np, pi, pj, s, cs, perm, offsets)
                                : hc_v1<float>(n, A, np, pi, pj, s, cs, perm, offsets);
    }
    return SPOPS_SUCCESS;
  });
}
