// convert_vnec.cpp -- CRS -> VNEC host builder.
//
// The reference's VNEC set-up (reference include/VNEC.h:78-227) computes
// only the variant selector; its per-thread arrays are never kept (locals
// shadow the members).  This builder produces the structure its kernels
// read (reference src/rvv/spmv_mtx_vnec.cpp:192-586):
//  * merge-path partition of rows + entries into NUM_THREADS_VNEC equal
//    diagonals, coordinates found by the search of MergePathDivide
//    (reference src/rvv/spmv_mtx_vnec.cpp:42-71; Merrill & Garland, SC 2016);
//  * per partition, the ascending list of columns its entries touch
//    (use_x_indices, NEC_NUM of them) and, per entry, the rank of its column
//    in that list (ecr_indices) -- the non-empty-column compaction of
//    include/VNEC.h:104-140;
//  * IRD = nnz / (lane groups * nLanes) with the grouping of
//    include/VNEC.h:144-178 (a new group starts when a compacted column is
//    nLanes or more past the group's first), partitions walked in order;
//    variant L for m >= 2,000,000, else D if IRD >= 0.477 (FP64) / 0.437
//    (FP32), else S (include/VNEC.h:216-226).
#include <omp.h>

#include <algorithm>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

// The merge-path coordinate of `diagonal`: a = row end offsets, b = entry
// indices 0..nnz-1 (b[k] = k).
static coord merge_path(int64_t diagonal, const int* row_end, int64_t m, int64_t nnz) {
  int64_t lo = std::max<int64_t>(diagonal - nnz, 0), hi = std::min<int64_t>(diagonal, m);
  while (lo < hi) {
    const int64_t pivot = (lo + hi) >> 1;
    if ((int64_t)row_end[pivot] <= diagonal - pivot - 1) lo = pivot + 1;
    else hi = pivot;
  }
  return coord{(int)std::min<int64_t>(lo, m), (int)(diagonal - lo)};
}

template <class V>
void build_vnec(spMtxVNEC<V>& d, const spMtxCRS<V>& s, int nparts) {
  const int64_t m = (int64_t)s.m, n = (int64_t)s.n, nnz = (int64_t)s.nz;
  if (nparts < 1) throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: at least one partition");
  d = spMtxVNEC<V>();
  d.G = &s;
  d.m = (int)m;
  d.n = (int)n;
  d.nnz = (int)nnz;
  d.NUM_THREADS_VNEC = nparts;
  const int64_t items = m + nnz, per = (items + nparts - 1) / nparts;
  d.thread_coord_start.resize((size_t)nparts);
  d.thread_coord_end.resize((size_t)nparts);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nparts; ++t) {
    const int64_t d0 = std::min<int64_t>(per * t, items), d1 = std::min<int64_t>(d0 + per, items);
    d.thread_coord_start[(size_t)t] = merge_path(d0, s.Rst + 1, m, nnz);
    d.thread_coord_end[(size_t)t] = merge_path(d1, s.Rst + 1, m, nnz);
  }
  // non-empty-column compaction per partition
  std::vector<std::vector<int>> cols((size_t)nparts);
  d.ecr_indices.assign((size_t)nnz, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nparts; ++t) {
    const int64_t y0 = d.thread_coord_start[(size_t)t].y, y1 = d.thread_coord_end[(size_t)t].y;
    std::vector<int>& c = cols[(size_t)t];
    c.assign(s.Col + y0, s.Col + y1);
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    for (int64_t e = y0; e < y1; ++e)
      d.ecr_indices[(size_t)e] = (int)(std::lower_bound(c.begin(), c.end(), s.Col[e]) - c.begin());
  }
  d.NEC_NUM.resize((size_t)nparts);
  d.nec_ptr.assign((size_t)nparts + 1, 0);
  for (int t = 0; t < nparts; ++t) {
    d.NEC_NUM[(size_t)t] = (int)cols[(size_t)t].size();
    d.nec_ptr[(size_t)t + 1] = d.nec_ptr[(size_t)t] + d.NEC_NUM[(size_t)t];
  }
  d.use_x_indices.resize((size_t)d.nec_ptr[(size_t)nparts]);
  for (int t = 0; t < nparts; ++t)
    std::copy(cols[(size_t)t].begin(), cols[(size_t)t].end(),
              d.use_x_indices.begin() + d.nec_ptr[(size_t)t]);
  vnec_select_variant(d, s);
}

// IRD: lane groups over the compacted columns, partitions in order; rows
// [start.x, end.x) of a partition, each from max(row start, start.y)
template <class V>
void vnec_select_variant(spMtxVNEC<V>& d, const spMtxCRS<V>& s) {
  const int64_t m = (int64_t)s.m, nnz = (int64_t)s.nz;
  const int nparts = d.NUM_THREADS_VNEC;
  const int lanes = std::is_same<V, float>::value ? nLanes_f32 : nLanes_f64;
  int64_t groups = 0;
  for (int t = 0; t < nparts; ++t) {
    const coord a = d.thread_coord_start[(size_t)t], b = d.thread_coord_end[(size_t)t];
    for (int64_t r = a.x; r < b.x; ++r) {
      const int64_t e0 = std::max<int64_t>(s.Rst[r], a.y), e1 = s.Rst[r + 1];
      if (e1 <= e0) continue;
      int64_t first = d.ecr_indices[(size_t)e0];
      for (int64_t e = e0 + 1; e < e1; ++e)
        if (d.ecr_indices[(size_t)e] - first >= lanes) {
          ++groups;
          first = d.ecr_indices[(size_t)e];
        }
      ++groups;
    }
  }
  d.IRD_mat = groups > 0 ? (float)nnz / (float)(groups * lanes) : 0.0f;
  const double thr = std::is_same<V, float>::value ? IRD_thr_fp32 : IRD_thr_fp64;
  if (m < 2000000) d.vnec_t = d.IRD_mat >= thr ? SpB_VNEC_D : SpB_VNEC_S;
  else d.vnec_t = SpB_VNEC_L;
}

template void vnec_select_variant<double>(spMtxVNEC<double>&, const spMtxCRS<double>&);
template void vnec_select_variant<float>(spMtxVNEC<float>&, const spMtxCRS<float>&);
template void build_vnec<double>(spMtxVNEC<double>&, const spMtxCRS<double>&, int);
template void build_vnec<float>(spMtxVNEC<float>&, const spMtxCRS<float>&, int);

}  // namespace rvvla
