// Device matrix object behind the opaque spops_matrix handle.
#pragma once

#include <vector>

#include "common.hpp"

namespace spops {

constexpr int kCsrStreamNnz = 512;     // entries staged per wave stream block
constexpr int64_t kCsrLongRow = 2048;  // longer rows: one workgroup each
constexpr int64_t kSpmmLongRow = 256;  // SpMM: longer rows are summed in segments
// SpMM: entries per segment of a long row (R-MAT 20, k = 64: forward 0.72 -> 0.68 ms against 512;
// 2048 no better)
constexpr int64_t kSpmmSegment = 1024;
constexpr int64_t kSpmmWindow = 4096;  // SpMM: short rows are ordered by length inside windows of this many rows
constexpr int64_t kSddmmChunk = 64;    // SDDMM by rows: entries of a row per lane-group task
constexpr int kBsrMaxBlock = 256;      // bR * bC limit of the wave-per-block-row VALU kernel
// MFMA BSR kernel: blocks with bR >= 8, bC >= 8, bC % 4 == 0, at most these sizes
constexpr int kBsrMfmaMinDim = 8, kBsrMfmaMaxRows = 64, kBsrMfmaMaxBlock = 4096;
constexpr int64_t kRowSegment = (int64_t)1 << 24;   // entries per buffer resource of a long row

// Column-sliced stream layout (SELL-C-sigma, SELL-C-R, ELL)
constexpr int kXcds = 8;                  // gfx950: workgroup g runs on XCD g % 8
constexpr int64_t kSliceBytesDefault = (int64_t)2 << 20;   // x bytes per slice (one XCD's L2 share)
constexpr int kSliceTaskRows = 256;       // rows of a multi-row wave task (row sums staged in LDS)
// entries per lane of a multi-row task: 16 (8 for FP64 values, which need twice the registers)
constexpr int slice_lane_entries(int vt) { return vt == SPOPS_VALUE_F64 ? 8 : 16; }

// A fixed permutation y[i] = x[g[i]] of n elements applied as two passes instead of a random
// gather (one cache line per element): pass 1 takes tiles of consecutive sources, orders each in
// LDS by destination tile (loc1: the element's place there) and writes it back to its range of
// a staging array (coalesced); in the staging array each destination tile's elements then form
// one run per source tile.  Pass 2 takes one destination tile per workgroup, reads its runs
// (src2: the staging addresses, run after run), places the elements in LDS at their
// destinations (loc2) and writes the tile out (coalesced).  Built on the host once per
// permutation, for one element size (tiles of 128 KiB).
constexpr int kPermTile = 16384;
constexpr int kPermPad = 96;     // staging elements between source tiles (breaks the power-of-two stride)
struct PermPlan {
  DevBuf loc1, src2, loc2;    // uint16 [n], int32 [n], uint16 [n]
  int64_t n = -1;             // -1: not built
  int64_t tile = 0;           // elements per tile (kPermTile for 8-byte elements, twice that otherwise)
};

struct CsrDev {
  DevBuf rowptr, col, val, tasks, longs;
  DevBuf tperm;                 // transposed view: entry e holds the source matrix's entry tperm[e]
  DevBuf sorder;                // SpMM: the rows of at most kSpmmLongRow entries, by length inside windows
  DevBuf slongs;                // SpMM: the longer rows
  DevBuf lsegs;                 // SpMM: segments of the long rows, int2 {index in slongs, entry offset}
  DevBuf lsegoff;               // SpMM: first segment of each long row (nslong + 1 ints)
  mutable DevBuf spmm_part;     // SpMM: partial sums of the segments (grown on demand)
  mutable DevBuf erow;          // SDDMM: row of every entry (built on first use)
  mutable DevBuf tdv;           // transposed view: the fused backward's dot products in its own entry order
  mutable DevBuf pstage;        // transposed view: staging of the two-pass permutations
  mutable PermPlan pin, pout;   // transposed view: source order -> its order (values), and back (built on first use)
  DevBuf sdtasks;               // SDDMM by rows: int2 {row, first entry in the row}, chunks of kSddmmChunk
  int64_t nsdtasks = 0;
  bool rowptr64 = false, tperm64 = false;
  int ntasks = 0, nlong = 0, nshort = 0, nslong = 0, nlseg = 0;
};

// The columns are cut into P slices of about equal entry count; slice s is served by the
// XCDs q with q % P == s, so each XCD gathers from one slice of x only and that slice stays
// in its L2.  Inside a slice every row contributes a segment (its entries in the slice,
// or one dummy entry with column -1 when it has none), stored packed in row order.  A
// wave task covers consecutive rows and up to 64 K entries (K = slice_lane_entries): lane
// l loads entries K l .. K l + K - 1 with 128-bit loads and sums the row pieces it holds;
// a segmented scan across lanes joins rows that span lanes; row sums are staged in LDS and
// written out coalesced (a task covers at most kSliceTaskRows rows).  Per task and lane one word
// tells which of the lane's entries start a row (low K bits) and how many rows start in
// earlier lanes (high 16 bits), so a row's index is r0 + that count + its rank in the
// lane.  Task entry ranges start 16-byte aligned (dummy entries pad the gaps).  Segments
// longer than 64 K entries are wave tasks of their own; rows longer than kCsrLongRow are
// workgroup tasks over all of x.  Each slice writes one double partial per row; a second
// pass sums the P partials (P == 1: tasks write y directly).
struct SliceDev {
  DevBuf col, val;            // entries, slice-major, row order inside a slice
  DevBuf meta;                // uint32 [task][64]: row-start bits | rows started before << 16
  DevBuf tasks;               // int4 {r0, e0 (inside the slice), ne, rows | s << 16 | kind << 24}
  DevBuf lptr, lcol, lval, lrow;   // rows longer than kCsrLongRow: one workgroup each
  DevBuf part;                // double [P][n_rows] (P > 1)
  int P = 1, nlong = 0, ntasks = 0;
  int64_t sbase[kXcds] = {0};
  int qoff[kXcds + 1] = {0};
  int64_t C = 0, sigma = 0;   // the reference parameters
  std::vector<int32_t> ref_reorder, ref_chunk_len;
};

// Swept layout (SELL-C-sigma, SELL-C-R and ELL of matrices whose rows gather from all over x):
// rows are cut into blocks of at most kSweepMaxRows rows (whose FP64 sums fill the 160 KiB of
// LDS of a CU); a workgroup adds the products of its block into LDS, and the block's entries
// are stored sorted by column, so the 64 lanes of a gather instruction read 64 consecutive
// entries of one sweep over x and fall on few cache lines.  An entry is one uint32 word,
// (column - cb * 2^(32 - rb)) << rb | row in the block (rb: bits of the largest block), plus
// its value; cb, the column block, comes from the entry's group.  Entries are packed in
// groups of kSweepGroup (one column block per group, the last group of a column block
// padded), lane-interleaved so that each 128-bit load of a wave reads one contiguous block
// and gather u of lane l takes the group's sorted entry 64 u + l.  A block's groups are cut
// into shares (contiguous group ranges), one workgroup each; a block of one share writes y,
// a block of more shares writes one partial plane per share (in the y type) and a second
// pass adds them in share order.  Two ways to form the blocks:
//  * one class: consecutive rows with about equal entry counts, H shares each (H from the CU
//    count), planes [H][n_rows];
//  * two classes (split length T > 0): the rows of at most T entries and the longer rows form
//    separate runs (each in row order), cut into blocks of about kSweepMaxRows rows.  A block
//    of long rows holds many entries per 128-byte line of x, so its sweep fetches each line
//    of x for many entries; such blocks get as many shares as their entry counts ask for
//    (planes per block, [shares][rows]), and a row map sends each position to its row.
// The second form cuts the gathered 64-byte pieces of x by ~40 % on a power-law matrix
// (n = 2^21, mean 15.5) but measured slower there (108 against 86 us): the shares' entry counts
// are less even than the one-class cut's, and blocks of one share sweep all of x, which one
// XCD's L2 does not hold.  So one class is the default.  (An in-kernel "last workgroup adds"
// handoff needs agent-scope release/acquire, which on this chip writes back and invalidates
// the XCD's L2 and with it the cached x: measured slower.)
constexpr int kSweepMaxRows = 20480;          // 20480 x 8 bytes = 160 KiB of LDS
constexpr int64_t kSweepRowsTypical = 16384;  // rows of a block for the layout choice
constexpr int kSweepBlock = 1024;             // threads per workgroup (one workgroup per CU)
constexpr int kSweepLaneEntries = 8;
constexpr int kSweepGroup = kWave * kSweepLaneEntries;
constexpr int kSweepTargetWorkgroups = 256;   // one per CU (default of g_sweep_workgroups)
constexpr int kSweepMaxSplit = 16;            // shares of a one-class block
constexpr int kSweepMaxShares = 64;           // shares of a two-class block

struct SweepDev {
  DevBuf word, val;           // per group: kSweepGroup words / values, lane-interleaved
  DevBuf src;                 // int32 per slot (CSR handles): its CSR entry, -1 for padding
  DevBuf gdesc;               // int per group: column block | valid entries << 16
  DevBuf wgrp;                // int per workgroup + 1: first group
  DevBuf wblk;                // int2 per workgroup: {row block, share}
  DevBuf binfo;               // int4 per row block: {first position, rows, shares, first plane element}
  DevBuf rowidx;              // int per position: its row (two classes only)
  DevBuf rtasks;              // int4 per task of the plane-adding pass (two classes only)
  DevBuf part;                // partial planes (y type)
  int rb = 0, H = 0, n_rb = 0, ngroups = 0, nwg = 0, nrtasks = 0;
  int64_t split = 0;          // T of the two-class form, 0 for one class
  int64_t plane_elems = 0;    // elements of all partial planes
};

struct BsrDev {
  DevBuf brp, bcol, data;
  int64_t mb = 0, bR = 0, bC = 0, nnzb = 0;
  bool mfma = false;          // the MFMA kernel (dense blocks of at least 8 x 8)
};

}  // namespace spops

struct spops_matrix_s {
  int format = 0, value_type = 0, index_type = 0;
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int64_t dev_bytes = 0, algo_bytes = 0, slots = 0, ref_nchunks = 0, ell_width = 0;
  spops::CsrDev csr;
  spops::SliceDev sell;
  spops::SweepDev sweep;      // SpMV layout used instead of `sell` (or beside `csr`) when sweep.H > 0
  spops::BsrDev bsr;
};

namespace spops {
extern int64_t g_slice_bytes;        // x bytes per slice for matrices created from now on
// SELL/ELL device layout of matrices created from now on: 0 = chosen from the x footprint,
// 1 = column-sliced stream, 2 = swept
extern int64_t g_sweep_mode;
// swept layout: workgroups aimed for per launch (H = this / fewest row blocks, 1..kSweepMaxSplit)
extern int64_t g_sweep_workgroups;
extern int64_t g_sweep_max_rows;     // swept layout: rows of the largest row block
// swept layout: 0 one class, T > 0 two classes split at row length T
extern int64_t g_sweep_split;
// BSR matrices created from now on with dense blocks of at least 8 x 8 (bC % 4 == 0) on the
// MFMA kernel: 0 never, 1 (default) when the VALU kernel cannot hold the block (more than
// kBsrMaxBlock entries), 2 always
extern int64_t g_bsr_mfma;
// the kernel a BSR matrix of bR x bC blocks gets (throws SPOPS_ERROR_LIMIT if none fits)
bool bsr_use_mfma(int64_t bR, int64_t bC);
// The reference's ave_nnz_each_row, Gini_coeff_row and nnz_closed2diag_ratio of a host CSR
// (src/sparse_features.cpp:609-760): the features the batched SpMM autotuner classifies by.
void row_features_host(const int64_t *rowptr, const int64_t *col, int64_t n_rows, int64_t n_cols, double out[3]);
extern int64_t g_dense_products;   // LeSpGEMM: products of a cluster from which it accumulates densely
extern int64_t g_task_products;    // LeSpGEMM: products per numeric task of a large cluster (0: per call)
extern int g_wide_cursors;         // LeSpGEMM: wide-cluster cursor variant (0 off, 1 on, 2 per call)
extern int64_t g_heavy_products;   // LeSpGEMM: products from which a cluster's bitmap is built in parts
// a CSR handle's swept SpMV copy of the values from its CSR values (stream-ordered)
void refresh_sweep_values(spops_matrix_s &A, hipStream_t s);
// Builds and uploads the CSR plan of a matrix: SpMV row tasks and long rows (> kCsrLongRow
// entries); for SpMM the short-row order and the segments of the rows > kSpmmLongRow.
void upload_csr_plan(CsrDev &P, const std::vector<int64_t> &rp, int64_t n_rows);
void build_csr_plan(const std::vector<int64_t> &rp, int64_t n_rows, std::vector<int2> &tasks, std::vector<int> &longs);
int64_t spmv_algo_bytes(const spops_matrix_s &M, int64_t sI);
void launch_spmv(const spops_matrix_s &M, double alpha, const void *x, double beta, void *y, hipStream_t s);
}
