/* VNEC.h -- merge-path partitions with per-partition non-empty-column
 * compaction (reference include/VNEC.h:35-324, include/SpB_mxv_VNEC.h:14-65,
 * kernels src/rvv/spmv_mtx_vnec.cpp:192-586).
 * The (rows + nnz) merge path is cut into NUM_THREADS_VNEC equal diagonals
 * (thread_coord_start/end give (row, nnz) coordinates).  In partition t the
 * columns it touches are ranked: use_x_indices[nec_ptr[t] .. nec_ptr[t+1])
 * lists them ascending (NEC_NUM[t] of them) and ecr_indices[j] is the rank
 * of entry j's column, so partition t reads x only through its compact
 * copy.  IRD_mat = nnz / (lane groups * nLanes) with nLanes 8 (FP64) / 16
 * (FP32) picks the variant as the reference does: L when m >= 2,000,000,
 * else D when IRD >= 0.477 (FP64) / 0.437 (FP32), else S.  On gfx950 the
 * compact copy is staged in LDS and ecr_indices travel as 16 bits. */
#pragma once
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

#define nLanes_f32 16
#define nLanes_f64 8
#define IRD_thr_fp32 0.437
#define IRD_thr_fp64 0.477

namespace SparseMatrixLib {

enum SpB_VNEC_type { SpB_VNEC_D, SpB_VNEC_S, SpB_VNEC_L };

struct coord {
  int x;
  int y;
};

template <typename ValT>
class spMtxVNEC {
 public:
  int NUM_THREADS_VNEC = 0;
  const spMtxCRS<ValT>* G = nullptr;  /* non-owning, as reference VNEC.h:56 */
  int m = 0, n = 0, nnz = 0;
  std::vector<coord> thread_coord_start;
  std::vector<coord> thread_coord_end;
  std::vector<int> ecr_indices;
  std::vector<int> NEC_NUM;
  std::vector<int> nec_ptr;
  std::vector<int> use_x_indices;
  float IRD_mat = 0.f;
  SpB_VNEC_type vnec_t = SpB_VNEC_L;
  mutable DeviceImage device;
};

}  // namespace SparseMatrixLib
