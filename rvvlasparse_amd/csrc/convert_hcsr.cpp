// convert_hcsr.cpp -- CRS -> HCSR host converter.
//
// Produces exactly the structure of the reference converter crs_to_HCSR
// (reference src/scalar/crs_to_hcsr.cpp:30-135): same block grid, the same
// kind rule (CSR-kind when rows * p < block nnz, compared in double,
// :70-76), the same prefix layout of block_ptr / row_offset (:79-86), the
// same in-block entry order (source CSR order, :107-119), the same local
// row_ptr contents and the kind flag in bits 31:30 (:130-134).  Errors map
// to RVVLA codes: R or C == 0 -> INVALID_ARG (:38), block-count overflow or
// a row_ptr longer than 2^30-1 -> LENGTH (:55,:87-90), nnz mismatch -> LOGIC
// (:91).  Unlike the reference, C > 65536 is rejected (col_idx is 16-bit).
#include <omp.h>

#include <algorithm>
#include <limits>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

template <class V>
void build_hcsr(spMtxHCSR<V>& d, const spMtxCRS<V>& s, size_t R, size_t C, double p) {
  if (R == 0 || C == 0) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: R and C must be positive");
  if (C > 65536) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: C must not exceed 65536 (16-bit col_idx)");
  d = spMtxHCSR<V>();
  d.R = R;
  d.C = C;
  d.m = s.m;
  d.n = s.n;
  d.nz = s.nz;
  d.b_m = (s.m + R - 1) / R;
  d.b_n = (s.n + C - 1) / C;
  const size_t bm = d.b_m, bn = d.b_n, m = s.m;
  const size_t nblocks = bm * bn;
  if (bn != 0 && nblocks / bn != bm) throw Error(RVVLA_ERR_LENGTH, "HCSR: too many blocks");

  // Pass 1: per-block entry counts and kinds; count / row_ptr length go to
  // slot b+1 so a prefix sum turns them into offsets.
  d.block_ptr.assign(nblocks + 1, 0);
  d.row_offset.assign(nblocks + 1, 0);
  std::vector<uint8_t> kind(nblocks, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (long long br = 0; br < (long long)bm; ++br) {
    const size_t r0 = (size_t)br * R, r1 = std::min(m, r0 + R);
    int* cnt = d.block_ptr.data() + (size_t)br * bn + 1;
    for (size_t e = (size_t)s.Rst[r0]; e < (size_t)s.Rst[r1]; ++e) ++cnt[(size_t)s.Col[e] / C];
    const double rows = (double)(r1 - r0);
    for (size_t j = 0; j < bn; ++j) {
      const size_t c = (size_t)cnt[j];
      uint8_t k = 0;
      uint32_t len = 0;
      if (c != 0) {
        if (rows * p < (double)c) { k = 1; len = (uint32_t)(r1 - r0 + 1); }
        else { k = 2; len = (uint32_t)c; }
      }
      kind[(size_t)br * bn + j] = k;
      d.row_offset[(size_t)br * bn + j + 1] = len;
    }
  }
  size_t nz_acc = 0, rp_acc = 0;
  for (size_t b = 1; b <= nblocks; ++b) {
    nz_acc += (size_t)d.block_ptr[b];
    rp_acc += (size_t)d.row_offset[b];
    d.block_ptr[b] = (int)nz_acc;
    d.row_offset[b] = (uint32_t)rp_acc;
  }
  if (rp_acc > ((size_t(1) << 30) - 1))
    throw Error(RVVLA_ERR_LENGTH, "HCSR: row_ptr has " + std::to_string(rp_acc) +
                                      " entries, more than the 30-bit row_offset field holds");
  if (nz_acc != s.nz) throw Error(RVVLA_ERR_LOGIC, "HCSR: nnz mismatch");

  // Pass 2: scatter entries block by block in source order.
  d.col_idx.resize(s.nz);
  d.values.resize(s.nz);
  d.row_ptr.assign(rp_acc, 0);
#pragma omp parallel
  {
    std::vector<size_t> fill(bn);
#pragma omp for schedule(dynamic, 1)
    for (long long br = 0; br < (long long)bm; ++br) {
      const size_t r0 = (size_t)br * R, r1 = std::min(m, r0 + R);
      const size_t blk0 = (size_t)br * bn;
      for (size_t j = 0; j < bn; ++j) fill[j] = (size_t)d.block_ptr[blk0 + j];
      for (size_t i = r0; i < r1; ++i) {
        const uint32_t li = (uint32_t)(i - r0);
        for (size_t e = (size_t)s.Rst[i]; e < (size_t)s.Rst[i + 1]; ++e) {
          const size_t c = (size_t)s.Col[e], j = c / C, blk = blk0 + j;
          const size_t pos = fill[j]++;
          d.col_idx[pos] = (uint16_t)(c - j * C);
          d.values[pos] = s.Val[e];
          const size_t ro = d.row_offset[blk];
          if (kind[blk] == 1) d.row_ptr[ro + li + 1] += 1;
          else d.row_ptr[ro + (pos - (size_t)d.block_ptr[blk])] = li;
        }
      }
      for (size_t j = 0; j < bn; ++j) {
        const size_t blk = blk0 + j;
        if (kind[blk] != 1) continue;
        uint32_t* rp = d.row_ptr.data() + d.row_offset[blk];
        for (size_t li = 0; li < r1 - r0; ++li) rp[li + 1] += rp[li];
      }
    }
  }
  for (size_t b = 0; b < nblocks; ++b) d.row_offset[b] |= (uint32_t)kind[b] << 30;
}

template void build_hcsr<double>(spMtxHCSR<double>&, const spMtxCRS<double>&, size_t, size_t, double);
template void build_hcsr<float>(spMtxHCSR<float>&, const spMtxCRS<float>&, size_t, size_t, double);

}  // namespace rvvla
