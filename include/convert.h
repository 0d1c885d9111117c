/* convert.h -- CRS -> format conversion (reference include/convert.h:21-31).
 * convert<T, Fmt, true> fills the host fields of `dist` and uploads the
 * device image once (dist.device); the same keys as the reference are read
 * (HCSR: "R", "C", "p"), plus each format's own tunables (see the format
 * headers).  Every non-CRS format also reads "device": 1 builds the layout
 * on the GPU from an uploaded copy of src (histograms, scans, sorts and
 * scatters), keeps that image as dist.device and fills the host fields
 * from it -- the same arrays as the host build. */
#pragma once
#include <map>
#include <string>
#include "CRS.h"

namespace SparseMatrixLib {

union convertValue {
  int i;
  double d;
};

struct convertParams {
  std::map<std::string, convertValue> param;
};

template <class T, template <class> class MtxDist, bool simd = true>
void convert(MtxDist<T>& dist, const spMtxCRS<T>& src, const convertParams& params);

}  // namespace SparseMatrixLib
