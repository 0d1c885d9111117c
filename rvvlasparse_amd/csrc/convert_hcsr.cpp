// convert_hcsr.cpp -- CRS -> HCSR host converter.
//
// Output contract: the arrays of the reference converter crs_to_HCSR
// (reference src/scalar/crs_to_hcsr.cpp:30-135) for the same R, C, p: the
// R x C block grid in row-major block order; block_ptr the first entry of
// every block; a block is CSR kind when rows * p < its entry count
// (compared in double, :70-76), COO kind when it has entries otherwise,
// empty when it has none; row_offset the start of the block's slice of
// row_ptr (rows + 1 local row pointers for CSR kind, one local row per
// entry for COO kind) with the kind in bits 31:30; entries inside a block
// in source CSR order; col_idx the column inside the block.  Errors: R or C
// == 0 -> INVALID_ARG, a block count or row_ptr length beyond the fields ->
// LENGTH, an entry count that does not add up -> LOGIC.  C > 65536 is
// rejected (col_idx is 16-bit).
//
// Method (the same as the device builder, k_convert_hcsr.hip): the CSR
// order is already block-row-major, so sorting each block-row's entries by
// the key (block column, source position) puts the whole matrix in block
// order with every block in CSR order.  The sorted keys give each block's
// entry count as a run length, a scan gives block_ptr, the kind rule and a
// second scan give row_offset, and one pass over the sorted entries writes
// col_idx, values and the row pointers.
#include <omp.h>

#include <algorithm>
#include <numeric>

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
  const int64_t bm = (int64_t)d.b_m, bn = (int64_t)d.b_n, m = (int64_t)s.m;
  if (bn != 0 && (size_t)(bm * bn) / (size_t)bn != (size_t)bm)
    throw Error(RVVLA_ERR_LENGTH, "HCSR: too many blocks");
  const int64_t nblocks = bm * bn;
  if (m > 0 && (int64_t)s.Rst[m] != (int64_t)s.nz)
    throw Error(RVVLA_ERR_LOGIC, "HCSR: Rst[m] does not equal nz");

  // 1. block order: per block-row, the entries sorted by (block column,
  //    position in the block-row); key = bc << 32 | position
  const int64_t nz = (int64_t)s.nz;
  std::vector<uint64_t> key((size_t)nz);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t br = 0; br < bm; ++br) {
    const int64_t e0 = s.Rst[br * (int64_t)R], e1 = s.Rst[std::min(m, (br + 1) * (int64_t)R)];
    for (int64_t e = e0; e < e1; ++e)
      key[(size_t)e] = ((uint64_t)((size_t)s.Col[e] / C) << 32) | (uint64_t)(e - e0);
    std::sort(key.begin() + e0, key.begin() + e1);
  }

  // 2. entry counts of the blocks (run lengths of the sorted block columns)
  //    and their scan: block_ptr
  std::vector<int64_t> cnt((size_t)nblocks, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t br = 0; br < bm; ++br) {
    const int64_t e0 = s.Rst[br * (int64_t)R], e1 = s.Rst[std::min(m, (br + 1) * (int64_t)R)];
    for (int64_t e = e0; e < e1;) {
      const uint64_t bc = key[(size_t)e] >> 32;
      int64_t f = e + 1;
      while (f < e1 && (key[(size_t)f] >> 32) == bc) ++f;
      cnt[(size_t)(br * bn + (int64_t)bc)] = f - e;
      e = f;
    }
  }
  d.block_ptr.assign((size_t)nblocks + 1, 0);
  std::vector<uint8_t> kind((size_t)nblocks, 0);
  std::vector<uint64_t> len((size_t)nblocks, 0);
  {
    int64_t acc = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
      d.block_ptr[(size_t)b] = (int)acc;
      acc += cnt[(size_t)b];
    }
    if (acc != nz) throw Error(RVVLA_ERR_LOGIC, "HCSR: block entry counts do not add up to nz");
    d.block_ptr[(size_t)nblocks] = (int)acc;
  }

  // 3. kinds (reference rule) and the row_ptr slice lengths; their scan is
  //    row_offset
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblocks; ++b) {
    const int64_t br = b / bn;
    const int64_t rows = std::min<int64_t>((int64_t)R, m - br * (int64_t)R);
    const int64_t c = cnt[(size_t)b];
    if (c == 0) continue;
    const bool csr = (double)rows * p < (double)c;
    kind[(size_t)b] = csr ? 1 : 2;
    len[(size_t)b] = csr ? (uint64_t)(rows + 1) : (uint64_t)c;
  }
  d.row_offset.assign((size_t)nblocks + 1, 0);
  uint64_t total = 0;
  for (int64_t b = 0; b < nblocks; ++b) {
    d.row_offset[(size_t)b] = (uint32_t)total;
    total += len[(size_t)b];
    if (total > (uint64_t(1) << 30) - 1)
      throw Error(RVVLA_ERR_LENGTH, "HCSR: row_ptr would exceed the 30-bit row_offset field");
  }
  d.row_offset[(size_t)nblocks] = (uint32_t)total;

  // 4. one pass over the sorted entries of each block-row: col_idx,
  //    values, and the local rows (COO kind) or local row pointers (CSR
  //    kind: the count of the block's entries in lower local rows, which
  //    ascend inside a block)
  d.col_idx.resize((size_t)nz);
  d.values.resize((size_t)nz);
  d.row_ptr.assign((size_t)total, 0);
#pragma omp parallel
  {
    std::vector<uint32_t> lrow;  // local row of each entry of the block-row
#pragma omp for schedule(dynamic, 1)
    for (int64_t br = 0; br < bm; ++br) {
      const int64_t r0 = br * (int64_t)R, r1 = std::min(m, r0 + (int64_t)R);
      const int64_t e0 = s.Rst[r0], e1 = s.Rst[r1];
      lrow.resize((size_t)(e1 - e0));
      for (int64_t r = r0; r < r1; ++r)
        for (int64_t e = s.Rst[r]; e < s.Rst[r + 1]; ++e) lrow[(size_t)(e - e0)] = (uint32_t)(r - r0);
      for (int64_t i = e0; i < e1;) {
        const int64_t bc = (int64_t)(key[(size_t)i] >> 32), b = br * bn + bc;
        const int64_t c = cnt[(size_t)b];
        const size_t ro = d.row_offset[(size_t)b];
        uint32_t* rp = d.row_ptr.data() + ro;
        uint32_t next_row = 0;  // CSR kind: local rows whose pointer is written
        for (int64_t k = 0; k < c; ++k) {
          const int64_t e = e0 + (int64_t)(key[(size_t)(i + k)] & 0xFFFFFFFFull);
          d.col_idx[(size_t)(i + k)] = (uint16_t)((size_t)s.Col[e] - (size_t)bc * C);
          d.values[(size_t)(i + k)] = s.Val[e];
          const uint32_t lr = lrow[(size_t)(e - e0)];
          if (kind[(size_t)b] == 2) {
            rp[k] = lr;
          } else {
            while (next_row <= lr) rp[next_row++] = (uint32_t)k;
          }
        }
        if (kind[(size_t)b] == 1)
          while (next_row <= (uint32_t)(r1 - r0)) rp[next_row++] = (uint32_t)c;
        i += c;
      }
    }
  }
  for (int64_t b = 0; b < nblocks; ++b) d.row_offset[(size_t)b] |= (uint32_t)kind[(size_t)b] << 30;
}

template void build_hcsr<double>(spMtxHCSR<double>&, const spMtxCRS<double>&, size_t, size_t, double);
template void build_hcsr<float>(spMtxHCSR<float>&, const spMtxCRS<float>&, size_t, size_t, double);

}  // namespace rvvla
