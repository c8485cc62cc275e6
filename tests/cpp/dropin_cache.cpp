// Host-side checks of the drop-in device cache (include/spops_dropin.h): an entry is reused
// only while the host matrix matches what the copy was made from.  The entries made here hold
// no device memory (the `make` callbacks only count), so nothing touches a GPU.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "SpOps.h"

static int failures = 0;
#define EXPECT(cond, what)                                 \
  do {                                                     \
    if (!(cond)) {                                         \
      std::printf("FAIL: %s (line %d)\n", what, __LINE__); \
      ++failures;                                          \
    }                                                      \
  } while (0)

int main() {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  int made = 0;
  auto make = [&](Entry &) { ++made; };
  const int n = 1000, nnz = 3000;
  // one set of host buffers stands for "freed and reallocated at the same address"
  int *rp = new_array<int>(n + 1);
  int *ci = new_array<int>(nnz);
  double *va = new_array<double>(nnz);
  auto fill = [&](int seed) {
    for (int r = 0; r <= n; ++r) rp[r] = 3 * r;
    for (int j = 0; j < nnz; ++j) {
      ci[j] = (j * 7 + seed) % n;
      va[j] = 0.5 + j + seed;
    }
  };
  CSR_Matrix<int, double> A;
  A.num_rows = n; A.num_cols = n; A.num_nnzs = nnz;
  A.row_offset = rp; A.col_index = ci; A.values = va;
  fill(0);
  auto get = [&]() { cached_entry(A.row_offset, csr_source(A), csr_digest(A), make); };

  for (int mode : {SPOPS_DROPIN_VERIFY_SAMPLED, SPOPS_DROPIN_VERIFY_FULL}) {
    state().entries.clear();
    state().verify = mode;
    made = 0;
    fill(0);
    get();
    EXPECT(made == 1, "first call builds the copy");
    get();
    EXPECT(made == 1, "unchanged matrix reuses the copy");
    fill(1);   // another matrix of the same shape at the same addresses
    get();
    EXPECT(made == 2, "different contents at the same address rebuild");
    A.num_cols = n + 1;
    get();
    EXPECT(made == 3, "different dimensions rebuild");
    A.num_cols = n;
    double *va2 = new_array<double>(nnz);
    for (int j = 0; j < nnz; ++j) va2[j] = va[j];
    A.values = va2;
    get();
    EXPECT(made == 4, "a different values array rebuilds");
    A.values = va;
    delete_array(va2);
    get();
    EXPECT(made == 5, "back to the first values array rebuilds");
    va[0] += 1.0;   // first entry: in every sample
    get();
    EXPECT(made == 6, "in-place edit of a sampled entry rebuilds");
    va[1] += 1.0;   // entry 1 is not among the 65 sampled of 3000
    get();
    EXPECT(made == (mode == SPOPS_DROPIN_VERIFY_FULL ? 7 : 6),
           "in-place edit of an unsampled entry: rebuilt in full mode only");
  }
  state().entries.clear();
  state().verify = SPOPS_DROPIN_VERIFY_OFF;
  made = 0;
  get();
  get();
  EXPECT(made == 2, "off mode rebuilds on every call");
  EXPECT(state().entries.empty(), "off mode caches nothing");
  state().transient.reset();
  delete_array(rp);
  delete_array(ci);
  delete_array(va);
  std::printf("%s\n", failures ? "dropin cache: FAILED" : "dropin cache: ok");
  return failures ? 1 : 0;
}
