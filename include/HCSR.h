/* HCSR.h -- hierarchical CSR (reference include/HCSR.h:39-114).
 * The matrix is tiled into R x C blocks in row-major block order
 * (b = br * b_n + bc).  block_ptr[b] is the first entry of block b;
 * row_offset[b] holds the block's start inside row_ptr in bits 29:0 and its
 * kind in bits 31:30 (0 empty, 1 local CSR, 2 COO); the final sentinel has
 * no kind.  A CSR-kind block owns rows+1 local offsets in row_ptr, a
 * COO-kind block one local row per entry; col_idx = column - bc * C. */
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

namespace SparseMatrixLib {

template <typename ValT>
class spMtxHCSR {
 public:
  size_t R = 0;
  size_t C = 0;
  size_t m = 0;
  size_t n = 0;
  size_t nz = 0;
  size_t b_m = 0;
  size_t b_n = 0;
  std::vector<int> block_ptr;
  std::vector<uint32_t> row_offset;
  std::vector<uint32_t> row_ptr;
  std::vector<uint16_t> col_idx;
  std::vector<ValT> values;
  mutable DeviceImage device;

  spMtxHCSR() = default;
  spMtxHCSR(size_t rows, size_t cols, size_t nnz, size_t r, size_t c)
      : R(r), C(c), m(rows), n(cols), nz(nnz), b_m((rows + r - 1) / r), b_n((cols + c - 1) / c) {}
};

}  // namespace SparseMatrixLib
