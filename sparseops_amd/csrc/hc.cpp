// Hierarchical clustering of candidate pairs into row clusters (host, sequential).
//
// Bit-exact restatement of the reference's merge process:
//   hierarchical_clustering_v0_fast  include/spgemm_utility.h:743-915  (version 0)
//   hierarchical_clustering_v1       include/spgemm_utility.h:609-733  (version 1)
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
#include <queue>
#include <unordered_map>

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

bool rows_sorted_unique(const Csr64 &A, int64_t n) {
  for (int64_t r = 0; r < n; ++r)
    for (int64_t j = A.rp[r] + 1; j < A.rp[r + 1]; ++j)
      if (A.col[j - 1] >= A.col[j]) return false;
  return true;
}

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
bool initial_has(int64_t np, const int64_t *pi, const int64_t *pj, int64_t i, int64_t j) {
  int64_t first = 0, len = np;
  while (len > 0) {
    int64_t half = len >> 1, mid = first + half;
    bool less = pi[mid] != i ? pi[mid] < i : pj[mid] < j;
    if (less) { first = mid + 1; len = len - half - 1; }
    else len = half;
  }
  return first < np && pi[first] == i && pj[first] == j;
}

template <typename V>
int64_t hc_v0_fast(int64_t n, const Csr64 &A, int64_t np, const int64_t *pi, const int64_t *pj, const V *score,
                   int64_t cs, int64_t *perm, int64_t *offsets) {
  std::vector<int64_t> cl(n), sz(n, 1);
  std::vector<char> valid(n, 1);
  for (int64_t i = 0; i < n; ++i) cl[i] = i;
  // Initial stream: group equal scores (counting sort over the distinct values), scores
  // descending, input position descending inside a score.
  std::vector<int64_t> keep;
  keep.reserve((size_t)np);
  for (int64_t t = 0; t < np; ++t) {
    int64_t a = std::min(pi[t], pj[t]), b = std::max(pi[t], pj[t]);
    if (a == b || a >= n || b >= n) continue;
    keep.push_back(t);
  }
  std::unordered_map<V, int64_t> bucket_of;
  std::vector<V> distinct;
  for (int64_t t : keep) {
    V s = score[t] == V(0) ? V(0) : score[t];
    if (bucket_of.emplace(s, 0).second) distinct.push_back(s);
  }
  std::sort(distinct.begin(), distinct.end(), [](V a, V b) { return a > b; });
  for (size_t q = 0; q < distinct.size(); ++q) bucket_of[distinct[q]] = (int64_t)q;
  std::vector<int64_t> bstart(distinct.size() + 1, 0);
  std::vector<int64_t> bid(keep.size());
  for (size_t q = 0; q < keep.size(); ++q) {
    V s = score[keep[q]] == V(0) ? V(0) : score[keep[q]];
    bid[q] = bucket_of[s];
    bstart[bid[q] + 1]++;
  }
  for (size_t q = 0; q < distinct.size(); ++q) bstart[q + 1] += bstart[q];
  std::vector<int64_t> order(keep.size());
  {
    std::vector<int64_t> fill(bstart.begin(), bstart.end() - 1);
    for (int64_t q = (int64_t)keep.size() - 1; q >= 0; --q) order[fill[bid[q]]++] = keep[q];   // descending position
  }
  std::priority_queue<Item<V>, std::vector<Item<V>>, ItemLess<V>> dyn;
  PairSet discovered((size_t)np / 8 + 1024);
  const bool su = rows_sorted_unique(A, n);
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
        if (!initial_has(np, pi, pj, a, b) && discovered.insert(a, b))
          dyn.push(Item<V>{(V)jaccard(A, a, b, su), a, b});
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

}  // namespace
}  // namespace spops

using namespace spops;

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
