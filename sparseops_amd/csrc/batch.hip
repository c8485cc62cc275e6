// Batched SpMM: many independent FP32 matrices, each with its own number of right-hand sides,
// in one persistent launch that a HIP graph can capture.  Per matrix an autotuner picks the
// format and the tile (no reference counterpart as an entry point: each matrix's product is
// SpOps_csr, src/spmv_csr.cpp:165-185, applied to k columns).
//
// Packed X and Y: matrix i's X block is n_cols x ld_i and its Y block n_rows x ld_i, row-major,
// at 16-byte aligned offsets; ld_i = k_i rounded up to the format's vector width (columns
// k_i .. ld_i - 1 are padding: Y's padding columns receive A times X's padding columns).
//
// Formats (both in the one kernel, chosen per task, wave-uniform):
//   kFmtChunk    : chunks of R = 64 / L rows (rows optionally ordered by length inside windows
//                  of sigma rows); a chunk's entries are stored column-major as {col, value}
//                  pairs (entry j of the chunk's row r at j * R + r, padding col = -1), each
//                  row's entries near the diagonal first and its far ones last; a wave streams
//                  its chunk through LDS in pieces of up to 4 KiB and the row's lanes read them there.  L lanes per
//                  row, each covering NV vectors of VEC consecutive columns (16-byte X and Y
//                  accesses for VEC = 4);
//   kFmtCsrGroup : CSR, a lane group of GS lanes per row (GS = 4..64, a power of two >= k),
//                  lanes over the k columns, eight row entries in flight (any k).
// Schedule: whole matrices go to the task queues of the 8 XCDs, largest first to the least
// loaded queue, so that a matrix's X (band and scattered entries alike) is gathered through
// one L2; task t of the launch is entry t / 8 of queue t % 8, and the persistent workgroup g
// (on XCD g % 8, the grid a multiple of 8) takes tasks g, g + grid, ... (no counter, so the
// launch is graph-safe).  Queues list matrices by decreasing cost, so the last tasks are the
// small ones.
//
// The autotuner (batch_autotune): the reference's matrix features (average row length, row
// Gini coefficient, share of entries near the diagonal; src/sparse_features.cpp:609-760)
// and k place each matrix in a class; for each class present the candidate plans are timed
// on the class's largest matrix, once, at creation, and the winner is used for the class.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "common.hpp"
#include "spmv_internal.hpp"

namespace spops {

constexpr int kFmtCsrGroup = 0, kFmtChunk = 1;
constexpr int kBatchTaskNnz = 4096;       // entries of a CSR task
constexpr int kBatchTaskEntries = 8192;   // stored entries of a chunk task
constexpr int kBatchMaxChunkK = 64;       // k of the chunk format (L <= 16 lanes of 4 columns)

struct BatchMat {
  int64_t rp;      // CSR: first rowptr entry; chunks: first entry of the row order
  int64_t e;       // CSR: first entry (col / val)
  int64_t x, y;    // X and Y offsets (floats, multiples of 4)
  int64_t cb;      // chunks: first chunk
  int32_t n_rows, n_cols, k, ld, fmt, lanes, vec, nv, sigma, pad;
};

namespace {

template <int VEC> struct FV { struct alignas(VEC * 4) V { float e[VEC]; }; };

constexpr int kPiece = 512;     // entries a wave stages in LDS at a time (4 KiB)

// One chunk of R = 64 / L rows: lane (r, q) accumulates row r's columns
// c0 + v VEC .. c0 + v VEC + VEC - 1 (v < NV, c0 = cb + q VEC NV).  The chunk's entries are
// streamed through the wave's LDS slot `stage` in pieces of at most kPiece entries (at most
// kPiece / R columns of the chunk), each piece fetched with 16-byte loads (up to 4 KiB in
// flight per wave),
// then read by the row's L lanes as a broadcast.
template <int VEC, int NV>
__device__ __forceinline__ void chunk_rows(const BatchMat &m, int64_t cidx, int chunk, const int2 *__restrict__ ent,
                                           const int64_t *__restrict__ cbase, const int *__restrict__ cwidth,
                                           const int *__restrict__ crow, const float *__restrict__ Xm,
                                           float *__restrict__ Ym, float alpha, float beta, int lane, int2 *stage) {
  using V = typename FV<VEC>::V;
  constexpr int U = VEC * NV >= 8 ? 4 : 8;            // entries per step (register budget)
  const int L = m.lanes, R = kWave / L, rl = lane / L, q = lane % L;
  const int width = __builtin_amdgcn_readfirstlane(cwidth[cidx]);
  const int slot = chunk * R + rl;
  const int row = slot < m.n_rows ? crow[m.rp + slot] : -1;
  const int ld = m.ld;
  // chunk columns per piece: at most kPiece / R, the pieces of a chunk of about equal size
  const int pmax = kPiece / R, npc = (width + pmax - 1) / max(pmax, 1);
  const int pj = npc > 1 ? (width + npc - 1) / npc : pmax;
  // buffer resources: loads past the chunk's entries return zeros, and a padding entry's
  // column -1 puts its X offset out of range, which loads zeros
  const rsrc_t re = make_rsrc(ent + cbase[cidx], (uint64_t)width * R * 8);
  const rsrc_t rx = make_rsrc(Xm, (uint64_t)m.n_cols * ld * 4);
  bu4 *st4 = reinterpret_cast<bu4 *>(stage);
  for (int cb = 0; cb < ld; cb += L * VEC * NV) {
    const int c0 = cb + q * VEC * NV;
    float acc[NV][VEC];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int e = 0; e < VEC; ++e) acc[v][e] = 0.f;
    for (int j0 = 0; j0 < width; j0 += pj) {
      const uint32_t boff = (uint32_t)(j0 * R) * 8;           // 16-byte aligned: R is even
      const int nld = (pj * R + 2 * kWave - 1) / (2 * kWave);  // 16-byte loads per lane (wave-uniform)
      bu4 t[kPiece / 2 / kWave];
#pragma unroll
      for (int i = 0; i < kPiece / 2 / kWave; ++i)
        if (i < nld) t[i] = bload<bu4>(re, boff + (uint32_t)(i * kWave + lane) * 16);
#pragma unroll
      for (int i = 0; i < kPiece / 2 / kWave; ++i)
        if (i < nld) st4[i * kWave + lane] = t[i];
      const int jn = min(pj, width - j0);
      for (int j = 0; j < jn; j += U) {
        int2 en[U];
#pragma unroll
        for (int u = 0; u < U; ++u) en[u] = j + u < jn ? stage[(j + u) * R + rl] : make_int2(-1, 0);
        V xv[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int v = 0; v < NV; ++v) xv[u][v] = bload<V>(rx, (uint32_t)((en[u].x * ld + c0 + v * VEC) * 4));
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const float a = __int_as_float(en[u].y);
#pragma unroll
          for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[v][e] = __builtin_fmaf(a, xv[u][v].e[e], acc[v][e]);
        }
      }
    }
    if (row >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (c0 + v * VEC >= ld) continue;
        V *dst = reinterpret_cast<V *>(Ym + (int64_t)row * ld + c0 + v * VEC);
        V o;
        if (beta != 0.f) {
          const V old = *dst;
#pragma unroll
          for (int e = 0; e < VEC; ++e) o.e[e] = alpha * acc[v][e] + beta * old.e[e];
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e) o.e[e] = alpha * acc[v][e];
        }
        *dst = o;
      }
    }
  }
}

// X rows staged in LDS per round of a workgroup (kBlock / kWave consecutive chunks, one per
// wave): the rows that the round's near-diagonal entries touch, when they fit.
constexpr int kWinFloats = 6144;   // 24 KiB

// One chunk of R = 64 / L rows whose row lanes cover all ld columns (L VEC NV >= ld), with X
// rows [wlo, wlo + wn) in the workgroup's LDS window `xw` (row stride ld): an entry whose
// column falls in the window reads X there, the others (the far entries, kept at the end of
// each row, and padding) gather from memory in steps that hold any.  Otherwise as chunk_rows.
template <int VEC, int NV>
__device__ __forceinline__ void chunk_rows_win(const BatchMat &m, int64_t cidx, int chunk, const int2 *__restrict__ ent,
                                               const int64_t *__restrict__ cbase, const int *__restrict__ cwidth,
                                               const int *__restrict__ crow, const float *__restrict__ Xm,
                                               float *__restrict__ Ym, float alpha, float beta, int lane, int2 *stage,
                                               const float *xw, int wlo, int wn) {
  using V = typename FV<VEC>::V;
  constexpr int U = VEC * NV >= 8 ? 4 : 8;            // entries per step (register budget)
  const int L = m.lanes, R = kWave / L, rl = lane / L, q = lane % L;
  const int width = __builtin_amdgcn_readfirstlane(cwidth[cidx]);
  const int slot = chunk * R + rl;
  const int row = slot < m.n_rows ? crow[m.rp + slot] : -1;
  const int ld = m.ld, c0 = q * VEC * NV;
  const int pmax = kPiece / R, npc = (width + pmax - 1) / max(pmax, 1);
  const int pj = npc > 1 ? (width + npc - 1) / npc : pmax;
  const rsrc_t re = make_rsrc(ent + cbase[cidx], (uint64_t)width * R * 8);
  const rsrc_t rx = make_rsrc(Xm, (uint64_t)m.n_cols * ld * 4);
  bu4 *st4 = reinterpret_cast<bu4 *>(stage);
  float acc[NV][VEC];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[v][e] = 0.f;
  for (int j0 = 0; j0 < width; j0 += pj) {
    const uint32_t boff = (uint32_t)(j0 * R) * 8;             // 16-byte aligned: R is even
    const int nld = (pj * R + 2 * kWave - 1) / (2 * kWave);    // 16-byte loads per lane (wave-uniform)
    bu4 t[kPiece / 2 / kWave];
#pragma unroll
    for (int i = 0; i < kPiece / 2 / kWave; ++i)
      if (i < nld) t[i] = bload<bu4>(re, boff + (uint32_t)(i * kWave + lane) * 16);
#pragma unroll
    for (int i = 0; i < kPiece / 2 / kWave; ++i)
      if (i < nld) st4[i * kWave + lane] = t[i];
    const int jn = min(pj, width - j0);
    for (int j = 0; j < jn; j += U) {
      int2 en[U];
#pragma unroll
      for (int u = 0; u < U; ++u) en[u] = j + u < jn ? stage[(j + u) * R + rl] : make_int2(-1, 0);
      V xv[U][NV];
      uint32_t d[U];
      bool far = false;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        d[u] = (uint32_t)(en[u].x - wlo);
        const bool in = d[u] < (uint32_t)wn;
        far |= !in;
        const uint32_t li = in ? d[u] : 0u;
#pragma unroll
        for (int v = 0; v < NV; ++v) xv[u][v] = *reinterpret_cast<const V *>(xw + li * ld + c0 + v * VEC);
      }
      if (__any(far)) {
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (d[u] >= (uint32_t)wn)
#pragma unroll
            for (int v = 0; v < NV; ++v) xv[u][v] = bload<V>(rx, (uint32_t)((en[u].x * ld + c0 + v * VEC) * 4));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float a = __int_as_float(en[u].y);
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int e = 0; e < VEC; ++e) acc[v][e] = __builtin_fmaf(a, xv[u][v].e[e], acc[v][e]);
      }
    }
  }
  if (row >= 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (c0 + v * VEC >= ld) continue;
      V *dst = reinterpret_cast<V *>(Ym + (int64_t)row * ld + c0 + v * VEC);
      V o;
      if (beta != 0.f) {
        const V old = *dst;
#pragma unroll
        for (int e = 0; e < VEC; ++e) o.e[e] = alpha * acc[v][e] + beta * old.e[e];
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) o.e[e] = alpha * acc[v][e];
      }
      *dst = o;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_batch_spmm(const BatchMat *__restrict__ mats, const int4 *__restrict__ tasks,
                                                       int ntasks, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col, const float *__restrict__ val,
                                                       const int2 *__restrict__ ent, const int64_t *__restrict__ cbase,
                                                       const int *__restrict__ cwidth, const int *__restrict__ crow,
                                                       const int2 *__restrict__ cwin,
                                                       const float *__restrict__ X, float *__restrict__ Y, float alpha,
                                                       float beta) {
  __shared__ int2 stage[kBlock / kWave][kPiece];
  __shared__ bu4 xwin4[kWinFloats / 4];
  const int lane = lane_id(), w = wave_id();
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const int4 tk = tasks[t];
    const int count = __builtin_amdgcn_readfirstlane(tk.z);
    if (count == 0) continue;                         // queue padding
    const int mi = __builtin_amdgcn_readfirstlane(tk.x);
    const int start = __builtin_amdgcn_readfirstlane(tk.y);
    const BatchMat m = mats[mi];
    const float *Xm = X + m.x;
    float *Ym = Y + m.y;
    if (m.fmt == kFmtCsrGroup) {
      const int k = m.k, ld = m.ld;
      const int GS = m.lanes, R = kWave / GS;
      const int g = lane / GS, c0 = lane % GS;
      const int *rp = rowptr + m.rp;
      const int64_t nnz = rp[m.n_rows];
      const rsrc_t rc = make_rsrc(col + m.e, (uint64_t)nnz * 4);
      const rsrc_t rv = make_rsrc(val + m.e, (uint64_t)nnz * 4);
      const rsrc_t rx = make_rsrc(Xm, (uint64_t)m.n_cols * ld * 4);
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
              const uint32_t o = p + u < p1 ? (uint32_t)(p + u) * 4 : kOOB;
              cc[u] = bload<int>(rc, o);
              vv[u] = bload<float>(rv, o);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[u] = bload<float>(rx, cv ? (uint32_t)((cc[u] * ld + c) * 4) : kOOB);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = __builtin_fmaf(vv[u], xv[u], acc);
          }
          if (cv) {
            float o = alpha * acc;
            if (beta != 0.f) o += beta * Ym[(int64_t)r * ld + c];
            Ym[(int64_t)r * ld + c] = o;
          }
        }
      }
    } else {
      const int sel = m.vec * 4 + m.nv;               // wave-uniform
      if (m.lanes * m.vec * m.nv >= m.ld) {
        // rounds of one chunk per wave, the X rows of the round's near entries in LDS
        const int end = start + count;
        const float *xw = reinterpret_cast<const float *>(xwin4);
        for (int base = start; base < end; base += kBlock / kWave) {
          int lo = INT32_MAX, hi = -1;
          for (int qq = 0; qq < kBlock / kWave && base + qq < end; ++qq) {
            const int2 b = cwin[m.cb + base + qq];
            if (b.x >= 0 && b.y >= b.x) {
              lo = min(lo, b.x);
              hi = max(hi, b.y);
            }
          }
          lo &= ~3;                                   // 16-byte aligned start (4 rows)
          int wn = hi >= lo ? hi - lo + 1 : 0;
          if ((int64_t)wn * m.ld > kWinFloats) wn = 0;
          wn = __builtin_amdgcn_readfirstlane(wn);
          lo = __builtin_amdgcn_readfirstlane(lo);
          __syncthreads();                            // the previous round's window reads are done
          if (wn > 0) {
            const rsrc_t rxw = make_rsrc(Xm, (uint64_t)m.n_cols * m.ld * 4);
            const int n4 = (wn * m.ld + 3) / 4;
            for (int i = (int)threadIdx.x; i < n4; i += kBlock)
              xwin4[i] = bload<bu4>(rxw, (uint32_t)lo * (uint32_t)m.ld * 4u + (uint32_t)i * 16u);
          }
          __syncthreads();
          const int ch = base + w;
          if (ch < end) {
            const int64_t cidx = m.cb + ch;
            switch (sel) {
              case 1 * 4 + 1: chunk_rows_win<1, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w], xw, lo, wn); break;
              case 2 * 4 + 1: chunk_rows_win<2, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w], xw, lo, wn); break;
              case 4 * 4 + 2: chunk_rows_win<4, 2>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w], xw, lo, wn); break;
              default: chunk_rows_win<4, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w], xw, lo, wn); break;
            }
          }
        }
        continue;
      }
      for (int ch = start + w; ch < start + count; ch += kBlock / kWave) {
        const int64_t cidx = m.cb + ch;
        switch (sel) {
          case 1 * 4 + 1: chunk_rows<1, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w]); break;
          case 2 * 4 + 1: chunk_rows<2, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w]); break;
          case 4 * 4 + 2: chunk_rows<4, 2>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w]); break;
          default: chunk_rows<4, 1>(m, cidx, ch, ent, cbase, cwidth, crow, Xm, Ym, alpha, beta, lane, stage[w]); break;
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
  int fmt = kFmtChunk, lanes = 1, vec = 4, nv = 1, sigma = 1;
};

struct BatchDev {
  DevBuf mats, tasks, rowptr, col, val, ent, cbase, cwidth, crow, cwin;
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

int pow2_at_least(int64_t v, int cap) {
  int p = 1;
  while (p < v && p < cap) p <<= 1;
  return p;
}

// the default plan of a matrix with k right-hand sides
MatPlan default_plan(int64_t k) {
  MatPlan p;
  if (k > kBatchMaxChunkK) {
    p.fmt = kFmtCsrGroup;
    p.lanes = pow2_at_least(k, kWave);
    return p;
  }
  p.fmt = kFmtChunk;
  p.vec = k == 1 ? 1 : k == 2 ? 2 : 4;
  p.nv = 1;
  p.lanes = pow2_at_least((k + p.vec - 1) / p.vec, 16);
  p.sigma = 1;
  return p;
}

int64_t plan_ld(const MatPlan &p, int64_t k) { return p.fmt == kFmtChunk ? (k + p.vec - 1) / p.vec * p.vec : k; }

// Device arrays for a set of matrices with the given plans.
void build_batch(const std::vector<HostMat> &ms, const std::vector<MatPlan> &plans, BatchDev &B) {
  const size_t nm = ms.size();
  std::vector<BatchMat> hm(nm);
  std::vector<int32_t> rowptr, col, cwidth, crow;
  std::vector<int2> cwin;                       // per chunk: column range of its near entries
  std::vector<float> val;
  std::vector<int2> ent;
  std::vector<int64_t> cbase;
  std::vector<std::vector<int4>> mq(nm);        // per matrix: its tasks in row order
  std::vector<int64_t> mcost(nm, 0);
  int64_t x = 0, y = 0, flop = 0, algo = 0;
  for (size_t i = 0; i < nm; ++i) {
    const HostMat &M = ms[i];
    const MatPlan &P = plans[i];
    BatchMat &b = hm[i];
    const int64_t ld = plan_ld(P, M.k);
    b.n_rows = (int32_t)M.n_rows;
    b.n_cols = (int32_t)M.n_cols;
    b.k = (int32_t)M.k;
    b.ld = (int32_t)ld;
    b.fmt = P.fmt;
    b.lanes = P.lanes;
    b.vec = P.vec;
    b.nv = P.nv;
    b.sigma = P.sigma;
    b.x = x;
    b.y = y;
    x += (M.n_cols * ld + 3) / 4 * 4;
    y += (M.n_rows * ld + 3) / 4 * 4;
    const int64_t nnz = M.rp[M.n_rows];
    SPOPS_REQUIRE(M.n_cols * ld * 4 < (int64_t)kMaxBufBytes && nnz * 4 < (int64_t)kMaxBufBytes, SPOPS_ERROR_LIMIT,
                  "matrix %zu: X block (n_cols * ld * 4 bytes) and entries (nnz * 4 bytes) must stay below 2 GiB", i);
    flop += 2 * nnz * M.k;
    algo += nnz * 8 + (M.n_rows + 1) * 4 + M.n_cols * M.k * 4 + M.n_rows * M.k * 4;
    mcost[i] = nnz * ((M.k + 3) / 4) + M.n_rows;
    if (P.fmt == kFmtCsrGroup) {
      b.rp = (int64_t)rowptr.size();
      b.e = (int64_t)col.size();
      for (int64_t r = 0; r <= M.n_rows; ++r) rowptr.push_back((int32_t)M.rp[r]);
      for (int64_t j = 0; j < nnz; ++j) { col.push_back((int32_t)M.ci[j]); val.push_back(M.v[j]); }
      int64_t r = 0;
      while (r < M.n_rows) {                   // runs of rows with about kBatchTaskNnz entries
        int64_t r1 = r, acc = 0;
        while (r1 < M.n_rows && (acc == 0 || acc + (M.rp[r1 + 1] - M.rp[r1] + 1) <= kBatchTaskNnz) && r1 - r < 4096)
          acc += M.rp[r1 + 1] - M.rp[r1] + 1, ++r1;
        mq[i].push_back(make_int4((int)i, (int)r, (int)(r1 - r), 0));
        r = r1;
      }
    } else {
      const int R = kWave / P.lanes;
      b.rp = (int64_t)crow.size();
      b.cb = (int64_t)cbase.size();
      std::vector<int64_t> order(M.n_rows);
      std::iota(order.begin(), order.end(), 0);
      if (P.sigma > 1)
        for (int64_t w0 = 0; w0 < M.n_rows; w0 += P.sigma) {
          const int64_t w1 = std::min<int64_t>(M.n_rows, w0 + P.sigma);
          std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int64_t a, int64_t c) {
            return M.rp[a + 1] - M.rp[a] > M.rp[c + 1] - M.rp[c];
          });
        }
      for (int64_t s = 0; s < M.n_rows; ++s) crow.push_back((int32_t)order[s]);
      // inside a row, the entries near the diagonal (|col - row scaled| <= the matrix's 90th
      // percentile distance) come first and the far ones last, so the steps of a chunk that
      // gather far X rows (L2 misses) are few and hold all of them
      std::vector<int64_t> dist(nnz);
      for (int64_t r = 0; r < M.n_rows; ++r)
        for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j)
          dist[j] = std::llabs(M.ci[j] - (M.n_rows ? r * M.n_cols / M.n_rows : 0));
      int64_t near = 0;
      if (nnz > 0) {
        std::vector<int64_t> d2(dist);
        std::nth_element(d2.begin(), d2.begin() + (nnz * 9) / 10, d2.end());
        near = d2[(nnz * 9) / 10];
      }
      std::vector<int64_t> eord(nnz);              // entry order inside each row
      for (int64_t r = 0; r < M.n_rows; ++r) {
        int64_t o = M.rp[r];
        for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j) if (dist[j] <= near) eord[o++] = j;
        for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j) if (dist[j] > near) eord[o++] = j;
      }
      const int64_t nch = (M.n_rows + R - 1) / R;
      std::vector<int64_t> ccost(nch);
      for (int64_t c = 0; c < nch; ++c) {
        int64_t wdt = 0;
        for (int64_t l = 0; l < R && c * R + l < M.n_rows; ++l) {
          const int64_t r = order[c * R + l];
          wdt = std::max(wdt, M.rp[r + 1] - M.rp[r]);
        }
        SPOPS_REQUIRE(wdt * R * 8 < (int64_t)kMaxBufBytes, SPOPS_ERROR_LIMIT, "matrix %zu: chunk of 2 GiB", i);
        int64_t wlo = INT32_MAX, whi = -1;
        for (int64_t l = 0; l < R && c * R + l < M.n_rows; ++l) {
          const int64_t r = order[c * R + l];
          for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j)
            if (dist[j] <= near) {
              wlo = std::min<int64_t>(wlo, M.ci[j]);
              whi = std::max<int64_t>(whi, M.ci[j]);
            }
        }
        cwin.push_back(whi >= 0 ? make_int2((int)wlo, (int)whi) : make_int2(-1, -1));
        cbase.push_back((int64_t)ent.size());
        cwidth.push_back((int32_t)wdt);
        ccost[c] = wdt * R + R;
        const size_t at = ent.size();
        ent.resize(at + (size_t)(wdt * R + 1) / 2 * 2, make_int2(-1, 0));   // chunks start 16-byte aligned
        for (int64_t l = 0; l < R && c * R + l < M.n_rows; ++l) {
          const int64_t r = order[c * R + l];
          for (int64_t j = M.rp[r]; j < M.rp[r + 1]; ++j) {
            const int64_t e = eord[j];
            float f = M.v[e];
            int bits;
            std::memcpy(&bits, &f, 4);
            ent[at + (size_t)(j - M.rp[r]) * R + l] = make_int2((int)M.ci[e], bits);
          }
        }
      }
      // a task: at least one chunk per wave of the workgroup (the waves split a task's chunks),
      // more while they fit kBatchTaskEntries
      int64_t c = 0;
      while (c < nch) {
        int64_t c1 = c, acc = 0;
        while (c1 < nch && (c1 - c < kBlock / kWave || acc + ccost[c1] <= kBatchTaskEntries)) acc += ccost[c1++];
        mq[i].push_back(make_int4((int)i, (int)c, (int)(c1 - c), 1));
        c = c1;
      }
    }
  }
  SPOPS_REQUIRE(col.size() < (size_t)INT32_MAX && ent.size() < ((size_t)1 << 40), SPOPS_ERROR_LIMIT,
                "a batch holds < 2^31 CSR entries");
  // queues: whole matrices to XCDs, largest first to the least loaded queue (so a matrix's X
  // is gathered through one L2), then interleaved t = 8 i + q
  std::vector<size_t> mo(nm);
  std::iota(mo.begin(), mo.end(), 0);
  std::stable_sort(mo.begin(), mo.end(), [&](size_t a, size_t c) { return mcost[a] > mcost[c]; });
  std::vector<std::vector<int4>> queue(kXcds);
  std::vector<int64_t> load(kXcds, 0);
  for (size_t i : mo) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    queue[q].insert(queue[q].end(), mq[i].begin(), mq[i].end());
    load[q] += mcost[i];
  }
  size_t qlen = 0;
  for (auto &qq : queue) qlen = std::max(qlen, qq.size());
  std::vector<int4> ht(qlen * kXcds, make_int4(0, 0, 0, 0));
  for (int q = 0; q < kXcds; ++q)
    for (size_t j = 0; j < queue[q].size(); ++j) ht[j * kXcds + q] = queue[q][j];
  SPOPS_REQUIRE(ht.size() < (size_t)INT32_MAX, SPOPS_ERROR_LIMIT, "too many batch tasks");
  B.mats.upload(hm);
  B.tasks.upload(ht);
  B.rowptr.upload(rowptr);
  B.col.upload(col);
  B.val.upload(val);
  B.ent.upload(ent);
  B.cbase.upload(cbase);
  B.cwidth.upload(cwidth);
  B.crow.upload(crow);
  B.cwin.upload(cwin);
  B.ntasks = (int)ht.size();
  // persistent grid: every workgroup resident at once (the task striding is static), a
  // multiple of 8 so that task t runs on XCD t % 8
  int dev = 0, cus = 0, per_cu = 0;
  SPOPS_HIP_CHECK(hipGetDevice(&dev));
  SPOPS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  SPOPS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_spmm, kBlock, 0));
  const int64_t want = std::min<int64_t>(B.ntasks, (int64_t)cus * std::max(per_cu, 1));
  B.grid = (int)std::max<int64_t>(kXcds, want / kXcds * kXcds);
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
                     B.ent.as<int2>(), B.cbase.as<int64_t>(), B.cwidth.as<int>(), B.crow.as<int>(), B.cwin.as<int2>(),
                     X, Y, alpha, beta);
  SPOPS_HIP_CHECK(hipGetLastError());
}

// matrix class from the reference's features (row_features_host): k bucket (1, 2, 3-4, 5-8,
// 9-16, 17-32, 33-64, > 64) x ave_nnz_each_row > 16 x Gini_coeff_row > 0.2 (irregular row
// lengths) x nnz_closed2diag_ratio > 0.5 (entries near the diagonal: X gathers stay local)
int matrix_class(const HostMat &M) {
  const int kb = M.k <= 1 ? 0 : M.k <= 2 ? 1 : M.k <= 4 ? 2 : M.k <= 8 ? 3 : M.k <= 16 ? 4 : M.k <= 32 ? 5
               : M.k <= 64 ? 6 : 7;
  if (M.n_rows == 0 || M.n_cols == 0) return kb * 8;
  double f[3];
  row_features_host(M.rp.data(), M.ci.data(), M.n_rows, M.n_cols, f);
  return kb * 8 + (f[0] > 16 ? 4 : 0) + (f[1] > 0.2 ? 2 : 0) + (f[2] > 0.5 ? 1 : 0);
}
constexpr int kClasses = 64;

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

// plan per matrix: tune = 0 -> default_plan; tune = 1 -> time the candidates on each class's
// largest matrix and use the winner for the class
std::vector<MatPlan> batch_autotune(const std::vector<HostMat> &ms, int tune, std::vector<int> &cls) {
  std::vector<MatPlan> plans(ms.size());
  cls.assign(ms.size(), 0);
  std::vector<int64_t> rep(kClasses, -1);
  auto nnz_of = [&](int64_t i) { return ms[i].rp[ms[i].n_rows]; };
  for (size_t i = 0; i < ms.size(); ++i) {
    cls[i] = matrix_class(ms[i]);
    if (rep[cls[i]] < 0 || nnz_of(i) > nnz_of(rep[cls[i]])) rep[cls[i]] = (int64_t)i;
    plans[i] = default_plan(ms[i].k);
  }
  if (!tune) return plans;
  std::vector<MatPlan> best(kClasses);
  std::vector<bool> have(kClasses, false);
  for (int c = 0; c < kClasses; ++c) {
    if (rep[c] < 0) continue;
    const HostMat &M = ms[rep[c]];
    if (M.rp[M.n_rows] == 0) continue;
    std::vector<MatPlan> cand;
    const MatPlan d = default_plan(M.k);
    cand.push_back(d);
    if (d.fmt == kFmtChunk) {
      MatPlan s = d;
      s.sigma = 1024;
      cand.push_back(s);
      if (d.vec == 4 && d.lanes >= 2) {          // two vectors per lane, half the lanes per row
        MatPlan w = d;
        w.nv = 2;
        w.lanes = d.lanes / 2;
        cand.push_back(w);
      }
      MatPlan g;
      g.fmt = kFmtCsrGroup;
      g.lanes = pow2_at_least(M.k, kWave);
      g.lanes = std::max(g.lanes, 4);
      cand.push_back(g);
    }
    double bt = 1e30;
    for (const MatPlan &p : cand) {
      const double t = time_plan(M, p);
      if (t < bt) { bt = t; best[c] = p; }
    }
    have[c] = true;
  }
  for (size_t i = 0; i < ms.size(); ++i) {
    if (!have[cls[i]]) continue;
    MatPlan p = best[cls[i]];
    // the class shares the plan's shape; lanes follow each matrix's own k
    const MatPlan d = default_plan(ms[i].k);
    if (p.fmt == kFmtChunk && d.fmt == kFmtChunk) {
      p.vec = d.vec;
      p.lanes = p.nv == 2 ? std::max(1, d.lanes / 2) : d.lanes;
      if (p.vec != 4) p.nv = 1, p.lanes = d.lanes;
    } else if (p.fmt == kFmtCsrGroup) {
      p.lanes = std::max(4, pow2_at_least(ms[i].k, kWave));
    } else {
      p = d;
    }
    plans[i] = p;
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

extern "C" int spops_batch_plan(spops_batch B, int64_t *x_off, int64_t *y_off, int32_t *ld, int32_t *fmt,
                                int32_t *lanes, int32_t *sigma, int32_t *cls) {
  return guarded([&]() {
    SPOPS_REQUIRE(B != nullptr, SPOPS_ERROR_INVALID_VALUE, "null batch");
    const auto &h = B->dev.hmats;
    for (size_t i = 0; i < h.size(); ++i) {
      if (x_off) x_off[i] = h[i].x;
      if (y_off) y_off[i] = h[i].y;
      if (ld) ld[i] = h[i].ld;
      if (fmt) fmt[i] = h[i].fmt;
      if (lanes) lanes[i] = h[i].lanes;
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
    SPOPS_REQUIRE((uintptr_t)d_X % 16 == 0 && (uintptr_t)d_Y % 16 == 0, SPOPS_ERROR_INVALID_VALUE,
                  "X and Y must be 16-byte aligned");
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
    SPOPS_REQUIRE((uintptr_t)d_X % 16 == 0 && (uintptr_t)d_Y % 16 == 0, SPOPS_ERROR_INVALID_VALUE,
                  "X and Y must be 16-byte aligned");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    SPOPS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      launch_batch(D, d_X, d_Y, alpha, beta, s);
    } catch (...) {                 // leave the stream out of capture mode
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
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
