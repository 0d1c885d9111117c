/* types.h -- status, operation, descriptor and stage types of the SpMV API.
 * Same names and enumerator order as reference include/types.h:18-40, so
 * code written against the reference compiles unchanged. */
#pragma once
#include <stdint.h>

namespace SparseMatrixLib {

/* code == 0 on success; otherwise one of the RVVLA_* codes of rvvla_hip.h. */
struct sparse_matrix_status {
  int code = 0;
};

enum sparse_operation_t {
  SPARSE_OPERATION_NON_TRANSPOSE,
  SPARSE_OPERATION_TRANSPOSE
};

struct sparse_matrix_descr {};

/* PREPARATION: stage b (and y) in device memory; MULT: the kernel;
 * GET_RESULTS: copy y back; ALL_STAGES: the three in order. */
enum sparse_matrix_mv_stage {
  SPARSE_MATRIX_MV_PREPARATION,
  SPARSE_MATRIX_MV_MULT,
  SPARSE_MATRIX_MV_GET_RESULTS,
  SPARSE_MATRIX_MV_ALL_STAGES
};

}  // namespace SparseMatrixLib
