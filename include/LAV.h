/* LAV.h -- column-segmented, row-sorted lane chunks plus a CSR remainder
 * (reference include/LAV.h:10-94, kernel semantics
 * src/scalar/spmv_mtx_lav.cpp:52-79).
 * Columns are ranked by population; the most populated columns holding a
 * (1 - T) share of the entries form the dense part, split by rank into
 * segments of segment_len columns.  Segment s stores values/columns at
 * segment_ptr[s]*SIMD_Lanes, lane-step major (entry = step*SIMD_Lanes +
 * lane).  Chunk c of segment s spans steps [chunk_offsets[(cc+1)*s + c],
 * chunk_offsets[(cc+1)*s + c + 1]); steps below chunk_offsets_full_mask[s*cc
 * + c] use all lanes, later steps the first mask[step] lanes.  Lane i of
 * chunk c in segment s accumulates into row out_order[(c + cc*s)*SIMD_Lanes
 * + i].  The other entries form sparse_part, multiplied first (with beta).
 * This library sorts rows once by their dense-part length (the same order
 * in every segment) so a wavefront can accumulate a chunk across all
 * segments in registers; lanes inside a masked step that have no entry
 * there carry a zero-valued pad. */
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

namespace SparseMatrixLib {

template <typename type>
struct spMtxLAV {
  size_t SIMD_Lanes = 64;
  int m = 0;
  int n = 0;
  int nz = 0;
  size_t segment_count = 0;
  size_t chunk_count = 0;
  double T = 0.2;
  int segment_len = 262144;

  std::vector<int> chunk_offsets;
  std::vector<int> chunk_offsets_full_mask;
  std::vector<size_t> segment_ptr;
  std::vector<uint32_t> out_order;
  std::vector<uint16_t> mask;
  std::vector<type> val;
  std::vector<uint32_t> col_id;
  spMtxCRS<type> sparse_part;
  mutable DeviceImage device;
};

}  // namespace SparseMatrixLib
