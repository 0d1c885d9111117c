/* spmv_mtx.h -- the multiply entry point (reference include/spmv_mtx.h:21-29).
 * simd = true selects the gfx950 HIP implementation of this library; the
 * specialisations for every format and for float/double live in
 * librvvla_hip.so (rvvlasparse_amd/csrc/cxx_api.cpp). */
#pragma once
#include <vector>
#include "types.h"

namespace SparseMatrixLib {

template <class T, template <class> class Mtx, bool simd = true,
          sparse_matrix_mv_stage = SPARSE_MATRIX_MV_ALL_STAGES>
sparse_matrix_status sparse_mv(sparse_operation_t type_op, T alpha, const Mtx<T>& mat,
                               sparse_matrix_descr descr, const std::vector<T>& b,
                               T beta, std::vector<T>& y);

}  // namespace SparseMatrixLib
