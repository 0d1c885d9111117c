/* CVR.h -- compressed vectorisation-oriented row streams
 * (reference include/CVR.h:22-81, kernel src/scalar/spmv_mtx_cvr.cpp:32-66).
 * Rows are fed to SIMD_LEN lanes; whenever a lane finishes a row it takes
 * the next row of its stream.  vPack_vec_vals/cols hold the packed entries
 * (pack k = entries k*SIMD_LEN .. +SIMD_LEN-1, lane = position % SIMD_LEN).
 * vPack_vec_record holds (pos, row) pairs in pos order plus a final
 * sentinel pair per stream: once the packs below pos/SIMD_LEN are consumed,
 * lane pos % SIMD_LEN is flushed into y[row] and restarts.  Stream t owns
 * packs [writeback_pos_record[t*(SIMD_LEN+2)], ...) and records from
 * writeback_pos_record[t*(SIMD_LEN+2)+1]; entries +2.. +SIMD_LEN+1 are the
 * rows still open in each lane when the stream ends.  With Nthreads == 1
 * this is exactly the single-stream layout the reference scalar kernel
 * walks. */
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

namespace SparseMatrixLib {

template <typename ValT>
class spMtxCVR {
 public:
  size_t nItems = 0;
  size_t numRows = 0;
  size_t numCols = 0;
  size_t SIMD_LEN = 64;
  int Nthreads = 1;
  int Nthreads_avail = 1;
  std::vector<ValT> vPack_vec_vals;
  std::vector<uint32_t> vPack_vec_cols;
  std::vector<int> vPack_vec_record;
  std::vector<int> writeback_pos_record;
  mutable DeviceImage device;
};

}  // namespace SparseMatrixLib
