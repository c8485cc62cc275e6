// Device matrix object behind the opaque spops_matrix handle.
#pragma once

#include <vector>

#include "common.hpp"

namespace spops {

constexpr int kCsrStreamNnz = 512;     // entries staged per wave stream block
constexpr int64_t kCsrLongRow = 2048;  // longer rows: one workgroup each
constexpr int64_t kSellOvfMin = 64;    // rows shorter than this never leave their chunk
constexpr int kBsrMaxBlock = 256;      // bR * bC limit of the wave-per-block-row kernel
constexpr int64_t kRowSegment = (int64_t)1 << 24;   // entries per buffer resource of a long row

struct CsrDev {
  DevBuf rowptr, col, val, tasks, longs;
  bool rowptr64 = false;
  int ntasks = 0, nlong = 0;
};

struct SellDev {
  DevBuf gcol, gval, rinfo, cbase, cwidth;   // chunked 4-entry groups
  DevBuf optr, ocol, oval, orow;             // overflow rows (CSR)
  int64_t nchunks = 0;
  int Cd = 32;                               // device chunk height
  int n_olong = 0, n_ovec = 0;
  int64_t C = 0, sigma = 0;                  // the reference parameters
  std::vector<int32_t> ref_reorder, ref_chunk_len;
};

struct BsrDev {
  DevBuf brp, bcol, data;
  int64_t mb = 0, bR = 0, bC = 0, nnzb = 0;
};

}  // namespace spops

struct spops_matrix_s {
  int format = 0, value_type = 0, index_type = 0;
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int64_t dev_bytes = 0, algo_bytes = 0, slots = 0, ref_nchunks = 0, ell_width = 0;
  spops::CsrDev csr;
  spops::SellDev sell;
  spops::BsrDev bsr;
};

namespace spops {
void launch_spmv(const spops_matrix_s &M, double alpha, const void *x, double beta, void *y, hipStream_t s);
}
