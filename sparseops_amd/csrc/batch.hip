// Batched SpMM: many independent FP32 matrices, each with its own number of right-hand sides,
// in one persistent launch that a HIP graph can capture.  Per matrix an autotuner picks the
// format and the tile (no reference counterpart as an entry point: each matrix's product is
// SpOps_csr, src/spmv_csr.cpp:165-185, applied to k columns).
//
// Formats (both in the one kernel, chosen per task, wave-uniform):
//   kFmtCsrGroup : CSR, a lane group of GS lanes per row (GS = 4..64, a power of two >= k),
//                  lanes over the k columns, eight row entries in flight;
//   kFmtSell     : SELL-C-sigma with C = 64 (one wave per chunk, one lane per row), chunk
//                  entries stored column-major (coalesced index / value loads), rows sorted by
//                  length inside windows of sigma rows (sigma = 64: no reordering), each lane
//                  keeping its row's k <= 32 outputs in registers.
// Tiles: a CSR task is a run of consecutive rows holding about kBatchTaskNnz entries; a SELL
// task is kBlock / kWave chunks.  Tasks are ordered by decreasing cost and the persistent
// workgroups take tasks g, g + grid, ... (no counter, so the launch is graph-safe).
//
// The autotuner (batch_autotune): features per matrix (rows, entries, mean / max row length,
// coefficient of variation of row lengths, k) place it in a class (k bucket x row-length
// regime); for each class present, both formats (and for SELL both sigma values) are timed
// on the class's largest matrix, once, at creation, and the winner is used for the class.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "common.hpp"
#include "spmv_internal.hpp"

namespace spops {

constexpr int kFmtCsrGroup = 0, kFmtSell = 1;
constexpr int kBatchMaxK = 32;            // right-hand sides of a SELL-format matrix
constexpr int kBatchTaskNnz = 4096;       // entries of a CSR task

struct BatchMat {
  int64_t rp;      // first rowptr entry (packed int32 row pointers, relative to e)
  int64_t e;       // first entry (packed col / val)
  int64_t x, y;    // X and Y offsets (floats)
  int64_t cb;      // first chunk (SELL)
  int32_t n_rows, n_cols, k, fmt, gs, sigma;
};

namespace {

__global__ __launch_bounds__(kBlock) void k_batch_spmm(const BatchMat *__restrict__ mats, const int4 *__restrict__ tasks,
                                                       int ntasks, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col, const float *__restrict__ val,
                                                       const int64_t *__restrict__ cbase, const int *__restrict__ cwidth,
                                                       const int *__restrict__ crow, const float *__restrict__ X,
                                                       float *__restrict__ Y, float alpha, float beta) {
  const int lane = lane_id(), w = wave_id();
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const int4 tk = tasks[t];
    const int mi = __builtin_amdgcn_readfirstlane(tk.x);
    const int start = __builtin_amdgcn_readfirstlane(tk.y), count = __builtin_amdgcn_readfirstlane(tk.z);
    const BatchMat m = mats[mi];
    const int k = m.k;
    const float *Xm = X + m.x;
    float *Ym = Y + m.y;
    if (m.fmt == kFmtCsrGroup) {
      const int GS = m.gs, R = kWave / GS;
      const int g = lane / GS, c0 = lane % GS;
      const int *rp = rowptr + m.rp;
      const int *cm = col + m.e;
      const float *vm = val + m.e;
      for (int r = start + w * R + g; r < start + count; r += (kBlock / kWave) * R) {
        const int p0 = rp[r], p1 = rp[r + 1];
        for (int cb = 0; cb < k; cb += GS) {
          const int c = cb + c0;
          const bool cv = c < k;
          float acc = 0.f;
          for (int p = p0; p < p1; p += 8) {
            int cc[8];
            float vv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const bool ok = p + u < p1;
              cc[u] = ok ? cm[p + u] : 0;
              vv[u] = ok ? vm[p + u] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[u] = (cv && p + u < p1) ? Xm[(int64_t)cc[u] * k + c] : 0.f;
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = __builtin_fmaf(vv[u], xv[u], acc);
          }
          if (cv) {
            float o = alpha * acc;
            if (beta != 0.f) o += beta * Ym[(int64_t)r * k + c];
            Ym[(int64_t)r * k + c] = o;
          }
        }
      }
    } else {
      // SELL-64: wave w takes chunks start + w, start + w + 4, ...
      for (int ch = start + w; ch < start + count; ch += kBlock / kWave) {
        const int64_t cidx = m.cb + ch;
        const int64_t base = cbase[cidx];
        const int width = cwidth[cidx];
        const int slot = ch * kWave + lane;                 // sorted row position
        const int row = slot < m.n_rows ? crow[m.rp + slot] : -1;
        float acc[kBatchMaxK];
#pragma unroll
        for (int c = 0; c < kBatchMaxK; ++c) acc[c] = 0.f;
        for (int j = 0; j < width; ++j) {
          const int cc = col[base + (int64_t)j * kWave + lane];
          const float v = val[base + (int64_t)j * kWave + lane];
          if (cc >= 0) {
            const float *xr = Xm + (int64_t)cc * k;
#pragma unroll
            for (int c = 0; c < kBatchMaxK; ++c)
              if (c < k) acc[c] = __builtin_fmaf(v, xr[c], acc[c]);
          }
        }
        if (row >= 0) {
#pragma unroll
          for (int c = 0; c < kBatchMaxK; ++c) {
            if (c < k) {
              float o = alpha * acc[c];
              if (beta != 0.f) o += beta * Ym[(int64_t)row * k + c];
              Ym[(int64_t)row * k + c] = o;
            }
          }
        }
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
struct HostMat {
  int64_t n_rows, n_cols, k;
  std::vector<int64_t> rp, ci;
  std::vector<float> v;
};

struct MatPlan {
  int fmt = kFmtCsrGroup, gs = 4, sigma = kWave;
};

struct BatchDev {
  DevBuf mats, tasks, rowptr, col, val, cbase, cwidth, crow;
  int ntasks = 0, grid = 0;
  int64_t x_len = 0, y_len = 0, flop = 0, algo_bytes = 0;
  std::vector<BatchMat> hmats;
  std::vector<MatPlan> plans;
  std::vector<int> cls;             // class of each matrix
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchDev() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

int gs_for(int64_t k) { return k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64; }

// Device arrays for a set of matrices with the given plans (tasks cost-ordered).
void build_batch(const std::vector<HostMat> &ms, const std::vector<MatPlan> &plans, BatchDev &B) {
  std::vector<BatchMat> hm(ms.size());
  std::vector<int32_t> rowptr, col;
  std::vector<float> val;
  std::vector<int64_t> cbase;
  std::vector<int32_t> cwidth, crow;
  struct T { int4 t; int64_t cost; };
  std::vector<T> tasks;
  int64_t x = 0, y = 0, flop = 0, algo = 0;
  for (size_t i = 0; i < ms.size(); ++i) {
    const HostMat &M = ms[i];
    const MatPlan &P = plans[i];
    BatchMat &b = hm[i];
    b.n_rows = (int32_t)M.n_rows;
    b.n_cols = (int32_t)M.n_cols;
    b.k = (int32_t)M.k;
    b.fmt = P.fmt;
    b.gs = P.gs;
    b.sigma = P.sigma;
    b.x = x;
    b.y = y;
    x += M.n_cols * M.k;
    y += M.n_rows * M.k;
    const int64_t nnz = M.rp[M.n_rows];
    flop += 2 * nnz * M.k;
    algo += nnz * 8 + (M.n_rows + 1) * 4 + M.n_cols * M.k * 4 + M.n_rows * M.k * 4;
    if (P.fmt == kFmtCsrGroup) {
      b.rp = (int64_t)rowptr.size();
      b.e = (int64_t)col.size();
      for (int64_t r = 0; r <= M.n_rows; ++r) rowptr.push_back((int32_t)M.rp[r]);
      for (int64_t j = 0; j < nnz; ++j) { col.push_back((int32_t)M.ci[j]); val.push_back(M.v[j]); }
      int64_t r = 0;
      while (r < M.n_rows) {                // runs of rows with about kBatchTaskNnz entries
        int64_t r1 = r, acc = 0;
        while (r1 < M.n_rows && (acc == 0 || acc + (M.rp[r1 + 1] - M.rp[r1]) <= kBatchTaskNnz) && r1 - r < 4096) {
          acc += M.rp[r1 + 1] - M.rp[r1];
          ++r1;
        }
        const int64_t cost = (acc + (r1 - r)) * ((M.k + P.gs - 1) / P.gs);
        tasks.push_back({make_int4((int)i, (int)r, (int)(r1 - r), 0), cost});
        r = r1;
      }
    } else {
      // SELL-64: rows sorted by length (descending, stable) inside windows of sigma rows
      b.rp = (int64_t)crow.size();
      b.cb = (int64_t)cbase.size();
      std::vector<int64_t> order(M.n_rows);
      std::iota(order.begin(), order.end(), 0);
      for (int64_t w0 = 0; w0 < M.n_rows; w0 += P.sigma) {
        const int64_t w1 = std::min<int64_t>(M.n_rows, w0 + P.sigma);
        std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int64_t a, int64_t c) {
          return M.rp[a + 1] - M.rp[a] > M.rp[c + 1] - M.rp[c];
        });
      }
      for (int64_t s = 0; s < M.n_rows; ++s) crow.push_back((int32_t)order[s]);
      const int64_t nch = (M.n_rows + kWave - 1) / kWave;
      for (int64_t c = 0; c < nch; ++c) {
        int64_t wdt = 0;
        for (int64_t l = 0; l < kWave && c * kWave + l < M.n_rows; ++l) {
          const int64_t r = order[c * kWave + l];
          wdt = std::max(wdt, M.rp[r + 1] - M.rp[r]);
        }
        cbase.push_back((int64_t)col.size());
        cwidth.push_back((int32_t)wdt);
        const size_t at = col.size();
        col.resize(at + (size_t)wdt * kWave, -1);
        val.resize(at + (size_t)wdt * kWave, 0.f);
        for (int64_t l = 0; l < kWave && c * kWave + l < M.n_rows; ++l) {
          const int64_t r = order[c * kWave + l];
          for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j) {
            col[at + (size_t)(j - M.rp[r]) * kWave + l] = (int32_t)M.ci[j];
            val[at + (size_t)(j - M.rp[r]) * kWave + l] = M.v[j];
          }
        }
      }
      const int per = kBlock / kWave;
      for (int64_t c = 0; c < nch; c += per) {
        const int cnt = (int)std::min<int64_t>(per, nch - c);
        int64_t cost = 0;
        for (int q = 0; q < cnt; ++q) cost += (int64_t)cwidth[b.cb + c + q] * M.k;
        tasks.push_back({make_int4((int)i, (int)c, cnt, 1), cost});
      }
    }
  }
  SPOPS_REQUIRE(col.size() < (size_t)INT32_MAX, SPOPS_ERROR_LIMIT, "a batch holds < 2^31 stored entries");
  std::stable_sort(tasks.begin(), tasks.end(), [](const T &a, const T &b) { return a.cost > b.cost; });
  std::vector<int4> ht(tasks.size());
  for (size_t i = 0; i < tasks.size(); ++i) ht[i] = tasks[i].t;
  B.mats.upload(hm);
  B.tasks.upload(ht);
  B.rowptr.upload(rowptr);
  B.col.upload(col);
  B.val.upload(val);
  B.cbase.upload(cbase);
  B.cwidth.upload(cwidth);
  B.crow.upload(crow);
  B.ntasks = (int)ht.size();
  int dev = 0, cus = 0;
  SPOPS_HIP_CHECK(hipGetDevice(&dev));
  SPOPS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  B.grid = (int)std::max<int64_t>(1, std::min<int64_t>(B.ntasks, (int64_t)cus * 8));
  B.x_len = x;
  B.y_len = y;
  B.flop = flop;
  B.algo_bytes = algo;
  B.hmats = hm;
}

void launch_batch(const BatchDev &B, const float *X, float *Y, float alpha, float beta, hipStream_t s) {
  if (B.ntasks == 0) return;
  hipLaunchKernelGGL(k_batch_spmm, dim3((unsigned)B.grid), dim3(kBlock), 0, s, B.mats.as<BatchMat>(),
                     B.tasks.as<int4>(), B.ntasks, B.rowptr.as<int>(), B.col.as<int>(), B.val.as<float>(),
                     B.cbase.as<int64_t>(), B.cwidth.as<int>(), B.crow.as<int>(), X, Y, alpha, beta);
  SPOPS_HIP_CHECK(hipGetLastError());
}

// matrix class: k bucket (1-2, 3-4, 5-8, 9-16, 17-32, > 32) x row-length regime (mean row
// length <= 16 / > 16) x irregularity (coefficient of variation of row lengths <= 0.5 / > 0.5)
int matrix_class(const HostMat &M) {
  const int kb = M.k <= 2 ? 0 : M.k <= 4 ? 1 : M.k <= 8 ? 2 : M.k <= 16 ? 3 : M.k <= 32 ? 4 : 5;
  double mean = 0, sq = 0;
  for (int64_t r = 0; r < M.n_rows; ++r) {
    const double l = (double)(M.rp[r + 1] - M.rp[r]);
    mean += l;
    sq += l * l;
  }
  mean /= std::max<int64_t>(M.n_rows, 1);
  const double var = sq / std::max<int64_t>(M.n_rows, 1) - mean * mean;
  const double cv = mean > 0 ? std::sqrt(std::max(var, 0.0)) / mean : 0.0;
  return kb * 4 + (mean > 16 ? 2 : 0) + (cv > 0.5 ? 1 : 0);
}

double time_plan(const HostMat &M, const MatPlan &P) {
  std::vector<HostMat> one{M};
  std::vector<MatPlan> pl{P};
  BatchDev B;
  build_batch(one, pl, B);
  DevBuf X, Y;
  X.alloc((size_t)std::max<int64_t>(B.x_len, 1) * 4);
  Y.alloc((size_t)std::max<int64_t>(B.y_len, 1) * 4);
  SPOPS_HIP_CHECK(hipMemset(X.p, 0, X.bytes));
  hipEvent_t a, b;
  SPOPS_HIP_CHECK(hipEventCreate(&a));
  SPOPS_HIP_CHECK(hipEventCreate(&b));
  launch_batch(B, X.as<float>(), Y.as<float>(), 1.f, 0.f, nullptr);
  SPOPS_HIP_CHECK(hipEventRecord(a, nullptr));
  constexpr int reps = 5;
  for (int i = 0; i < reps; ++i) launch_batch(B, X.as<float>(), Y.as<float>(), 1.f, 0.f, nullptr);
  SPOPS_HIP_CHECK(hipEventRecord(b, nullptr));
  SPOPS_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0;
  SPOPS_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// plan per matrix: tune = 0 -> the cost rule alone (CSR groups, SELL for k <= 2);
// tune = 1 -> time the candidates on each class's largest matrix and use the winner
std::vector<MatPlan> batch_autotune(const std::vector<HostMat> &ms, int tune, std::vector<int> &cls) {
  std::vector<MatPlan> plans(ms.size());
  cls.assign(ms.size(), 0);
  std::vector<int64_t> rep(64, -1);
  for (size_t i = 0; i < ms.size(); ++i) {
    cls[i] = matrix_class(ms[i]);
    const int64_t nnz = ms[i].rp[ms[i].n_rows];
    if (rep[cls[i]] < 0 || nnz > ms[rep[cls[i]]].rp[ms[rep[cls[i]]].n_rows]) rep[cls[i]] = (int64_t)i;
    MatPlan &p = plans[i];
    p.gs = gs_for(ms[i].k);
    p.fmt = (ms[i].k <= 2) ? kFmtSell : kFmtCsrGroup;
    p.sigma = kWave;
  }
  if (!tune) return plans;
  std::vector<MatPlan> best(64);
  for (int c = 0; c < 64; ++c) {
    if (rep[c] < 0) continue;
    const HostMat &M = ms[rep[c]];
    std::vector<MatPlan> cand;
    MatPlan csr;
    csr.fmt = kFmtCsrGroup;
    csr.gs = gs_for(M.k);
    cand.push_back(csr);
    if (M.k > 4) {                          // a wider lane group than k needs
      MatPlan c2 = csr;
      c2.gs = std::min(64, csr.gs * 2);
      cand.push_back(c2);
    }
    if (M.k <= kBatchMaxK) {
      MatPlan s1;
      s1.fmt = kFmtSell;
      s1.gs = gs_for(M.k);
      s1.sigma = kWave;
      cand.push_back(s1);
      MatPlan s2 = s1;
      s2.sigma = 1024;
      cand.push_back(s2);
    }
    double bt = 1e30;
    for (const MatPlan &p : cand) {
      const double t = time_plan(M, p);
      if (t < bt) { bt = t; best[c] = p; }
    }
  }
  for (size_t i = 0; i < ms.size(); ++i) {
    plans[i] = best[cls[i]];
    if (plans[i].fmt == kFmtCsrGroup) plans[i].gs = std::max(plans[i].gs, gs_for(ms[i].k));
  }
  return plans;
}

}  // namespace

}  // namespace spops

struct spops_batch_s {
  spops::BatchDev dev;
};

using namespace spops;

extern "C" int spops_batch_create(spops_batch *out, int64_t nmat, const int64_t *n_rows, const int64_t *n_cols,
                                  const int64_t *k, int index_type, const void *const *rowptr, const void *const *col,
                                  const void *const *values, int in_value_type, int tune) {
  return guarded([&]() {
    SPOPS_REQUIRE(out != nullptr && nmat >= 0, SPOPS_ERROR_INVALID_VALUE, "null output or negative count");
    SPOPS_REQUIRE(nmat == 0 || (n_rows && n_cols && k && rowptr && col && values), SPOPS_ERROR_INVALID_VALUE,
                  "null argument array");
    SPOPS_REQUIRE(index_type == SPOPS_INDEX_I32 || index_type == SPOPS_INDEX_I64, SPOPS_ERROR_INVALID_VALUE,
                  "index_type must be I32 or I64");
    SPOPS_REQUIRE(in_value_type == SPOPS_VALUE_F64 || in_value_type == SPOPS_VALUE_F32, SPOPS_ERROR_INVALID_VALUE,
                  "input values must be F64 or F32");
    std::vector<HostMat> ms((size_t)nmat);
    int64_t total_nnz = 0;
    for (int64_t i = 0; i < nmat; ++i) {
      HostMat &M = ms[i];
      M.n_rows = n_rows[i];
      M.n_cols = n_cols[i];
      M.k = k[i];
      SPOPS_REQUIRE(M.n_rows >= 0 && M.n_cols >= 0 && M.n_rows < (int64_t)INT32_MAX && M.n_cols < (int64_t)INT32_MAX,
                    SPOPS_ERROR_INVALID_VALUE, "matrix %lld: bad shape", (long long)i);
      SPOPS_REQUIRE(M.k >= 1 && M.k <= 4096, SPOPS_ERROR_INVALID_VALUE, "matrix %lld: k must be in [1, 4096]",
                    (long long)i);
      M.rp.resize(M.n_rows + 1);
      for (int64_t r = 0; r <= M.n_rows; ++r)
        M.rp[r] = index_type == SPOPS_INDEX_I32 ? static_cast<const int32_t *>(rowptr[i])[r]
                                                : static_cast<const int64_t *>(rowptr[i])[r];
      SPOPS_REQUIRE(M.rp[0] == 0, SPOPS_ERROR_INVALID_VALUE, "matrix %lld: rowptr must start at 0", (long long)i);
      for (int64_t r = 0; r < M.n_rows; ++r)
        SPOPS_REQUIRE(M.rp[r + 1] >= M.rp[r], SPOPS_ERROR_INVALID_VALUE, "matrix %lld: rowptr not monotone",
                      (long long)i);
      const int64_t nnz = M.rp[M.n_rows];
      M.ci.resize(nnz);
      M.v.resize(nnz);
      for (int64_t j = 0; j < nnz; ++j) {
        M.ci[j] = index_type == SPOPS_INDEX_I32 ? static_cast<const int32_t *>(col[i])[j]
                                                : static_cast<const int64_t *>(col[i])[j];
        SPOPS_REQUIRE(M.ci[j] >= 0 && M.ci[j] < M.n_cols, SPOPS_ERROR_INVALID_VALUE,
                      "matrix %lld: column out of range", (long long)i);
        M.v[j] = in_value_type == SPOPS_VALUE_F64 ? (float)static_cast<const double *>(values[i])[j]
                                                  : static_cast<const float *>(values[i])[j];
      }
      total_nnz += nnz;
    }
    SPOPS_REQUIRE(total_nnz < (int64_t)INT32_MAX, SPOPS_ERROR_LIMIT, "a batch holds < 2^31 entries");
    auto b = std::make_unique<spops_batch_s>();
    std::vector<MatPlan> plans = batch_autotune(ms, tune, b->dev.cls);
    build_batch(ms, plans, b->dev);
    b->dev.plans = plans;
    *out = b.release();
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_info(spops_batch B, int64_t *x_len, int64_t *y_len, int64_t *flop, int64_t *algo_bytes,
                                int64_t *ntasks) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr, SPOPS_ERROR_INVALID_VALUE, "null batch");
    if (x_len) *x_len = B->dev.x_len;
    if (y_len) *y_len = B->dev.y_len;
    if (flop) *flop = B->dev.flop;
    if (algo_bytes) *algo_bytes = B->dev.algo_bytes;
    if (ntasks) *ntasks = B->dev.ntasks;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_plan(spops_batch B, int64_t *x_off, int64_t *y_off, int32_t *fmt, int32_t *gs,
                                int32_t *sigma, int32_t *cls) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr, SPOPS_ERROR_INVALID_VALUE, "null batch");
    const auto &h = B->dev.hmats;
    for (size_t i = 0; i < h.size(); ++i) {
      if (x_off) x_off[i] = h[i].x;
      if (y_off) y_off[i] = h[i].y;
      if (fmt) fmt[i] = h[i].fmt;
      if (gs) gs[i] = h[i].gs;
      if (sigma) sigma[i] = h[i].sigma;
      if (cls) cls[i] = B->dev.cls[i];
    }
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_spmm(spops_batch B, float alpha, const float *d_X, float beta, float *d_Y, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr, SPOPS_ERROR_INVALID_VALUE, "null batch");
    SPOPS_REQUIRE((d_X || B->dev.x_len == 0) && (d_Y || B->dev.y_len == 0), SPOPS_ERROR_INVALID_VALUE, "null X or Y");
    launch_batch(B->dev, d_X, d_Y, alpha, beta, static_cast<hipStream_t>(stream));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_capture(spops_batch B, float alpha, const float *d_X, float beta, float *d_Y, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr, SPOPS_ERROR_INVALID_VALUE, "null batch");
    SPOPS_REQUIRE(stream != nullptr, SPOPS_ERROR_INVALID_VALUE, "capture needs a non-default stream");
    auto &D = B->dev;
    if (D.exec) { SPOPS_HIP_CHECK(hipGraphExecDestroy(D.exec)); D.exec = nullptr; }
    if (D.graph) { SPOPS_HIP_CHECK(hipGraphDestroy(D.graph)); D.graph = nullptr; }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    SPOPS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_batch(D, d_X, d_Y, alpha, beta, s);
    SPOPS_HIP_CHECK(hipStreamEndCapture(s, &D.graph));
    SPOPS_HIP_CHECK(hipGraphInstantiate(&D.exec, D.graph, nullptr, nullptr, 0));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_replay(spops_batch B, void *stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr && B->dev.exec != nullptr, SPOPS_ERROR_INVALID_VALUE, "no captured graph");
    SPOPS_HIP_CHECK(hipGraphLaunch(B->dev.exec, static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_batch_destroy(spops_batch B) {
  delete B;
  return SPOPS_SUCCESS;
}
