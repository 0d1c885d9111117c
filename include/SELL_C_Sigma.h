/* SELL_C_Sigma.h -- sliced ELLPACK with sigma-window row sorting
 * (reference include/SELL_C_Sigma.h:23-101).
 * Slice i covers sorted rows SBs[i] .. SBs[i]+Bs[i]-1; its width Cl[i] is
 * its longest row; entry (slot j, row k) sits at Cs[i] + j*Bs[i] + k
 * (column-major inside the slice, so lane k of a 64-wide wavefront reads
 * consecutive addresses); padding has value 0 and a valid column.
 * vPerm[sorted row] = original row.  This library builds C = 64. */
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

namespace SparseMatrixLib {

template <typename ValT>
class spMtxSELL_C_Sigma {
 public:
  size_t m = 0;
  size_t n = 0;
  size_t cnt_b = 0;
  size_t nz = 0;
  int C = 0;
  int Sigma = 0;
  std::vector<int> vCl;
  std::vector<int> vCs;
  std::vector<int> vBs;
  std::vector<int> vSBs;
  std::vector<uint32_t> vCol;
  std::vector<uint32_t> vPerm;
  std::vector<ValT> vVal;
  std::vector<int> vPartition;
  mutable DeviceImage device;

  void freeMem() {
    vCl.clear(); vCs.clear(); vBs.clear(); vSBs.clear();
    vCol.clear(); vPerm.clear(); vVal.clear(); vPartition.clear();
    vCl.shrink_to_fit(); vCs.shrink_to_fit(); vBs.shrink_to_fit(); vSBs.shrink_to_fit();
    vCol.shrink_to_fit(); vPerm.shrink_to_fit(); vVal.shrink_to_fit(); vPartition.shrink_to_fit();
  }
};

}  // namespace SparseMatrixLib
