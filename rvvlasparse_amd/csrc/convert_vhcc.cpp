// convert_vhcc.cpp -- CRS -> VHCC host builder.
//
// The reference ships no VHCC builder (its container only has setters,
// reference include/VHCC.h:80-112); the layout is the one its vector kernel
// reads (reference src/rvv/spmv_vhcc.cpp:73-162), after Tang et al.
// (CGO 2015):
//  * the CSR-ordered entries, zero-padded to whole vectors of vec_size;
//  * the vectors are split into num_threads partitions of (nearly) equal
//    vector counts: _thr_info[t].start_vec / end_vec;
//  * a partition whose first entry continues a row started earlier sends
//    that row's partial sum to its own overflow row m + t; a row still open
//    at a partition's last entry gets a forced row end there, written to the
//    real row; after all partitions, result[last_row] += result[overflow_row]
//    (with last_row == overflow_row == m + t when nothing continues);
//  * _veceor_ptr lists the vectors holding a row end (real or forced), then
//    a sentinel (the vector count);
//  * per such vector, n_masks scan masks (lane j adds lane j - 2^s in step
//    s iff both lie in the same segment) and one write mask (the row-end
//    lanes), vec_size/8 bytes each, bit j of byte j/8 = lane j;
//  * _row_arr lists the target row of every row end in stream order.
#include <omp.h>

#include <algorithm>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

template <class V>
void build_vhcc(spMtxVHCC<V>& d, const spMtxCRS<V>& s, int vec, int nparts) {
  if (vec < 8 || vec > 64 || (vec & (vec - 1)))
    throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: vec_size must be 8, 16, 32 or 64");
  const int64_t m = (int64_t)s.m, nnz = (int64_t)s.nz;
  const int64_t nvec = (nnz + vec - 1) / vec;
  if (nparts < 1) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: at least one partition");
  nparts = (int)std::max<int64_t>(1, std::min<int64_t>(nparts, std::max<int64_t>(nvec, 1)));
  const int nm = 0 == vec ? 0 : [](int v) { int b = 0; while ((1 << b) < v) ++b; return b; }(vec);
  const int mb = vec / 8;  // bytes per mask
  d.freeMem();
  d.set_params(vhcc_params((size_t)nparts, 1, 512, 512, (size_t)vec));
  d.set_properties((size_t)m, s.n, (size_t)nnz, (size_t)nvec, (size_t)(nvec * vec - nnz),
                   (size_t)(m + nparts));
  auto& vals = d.get_vals_arr_reference();
  auto& cols = d.get_col_arr_reference();
  vals.assign((size_t)(nvec * vec), V(0));
  cols.assign((size_t)(nvec * vec), 0u);
  for (int64_t e = 0; e < nnz; ++e) {
    vals[(size_t)e] = s.Val[e];
    cols[(size_t)e] = (uint32_t)s.Col[e];
  }
  // row of every entry, and whether it ends its row
  std::vector<int32_t> row_of((size_t)nnz);
  std::vector<uint8_t> is_end((size_t)nnz, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; ++r)
    for (int e = s.Rst[r]; e < s.Rst[r + 1]; ++e) {
      row_of[(size_t)e] = (int32_t)r;
      is_end[(size_t)e] = e == s.Rst[r + 1] - 1;
    }
  // partitions and forced ends
  auto& thr = d.get_thr_info_reference();
  thr.assign((size_t)nparts, thr_info_t{});
  std::vector<int64_t> target((size_t)std::max<int64_t>(nnz, 1), -1);  // row_arr target per end
  for (int t = 0; t < nparts; ++t) {
    thr_info_t& T = thr[(size_t)t];
    T.start_vec = (int)(nvec * t / nparts);
    T.end_vec = (int)(nvec * (t + 1) / nparts);
    T.panel_id = 0;
    T.merge_start = (int)std::min<int64_t>((int64_t)T.start_vec * vec, nnz);
    T.merge_end = (int)std::min<int64_t>((int64_t)T.end_vec * vec, nnz);
    T.last_row = T.overflow_row = (int)(m + t);
  }
  for (int64_t e = 0; e < nnz; ++e)
    if (is_end[(size_t)e]) target[(size_t)e] = row_of[(size_t)e];
  for (int t = 0; t < nparts; ++t) {
    thr_info_t& T = thr[(size_t)t];
    const int64_t a = T.merge_start, b = T.merge_end;  // entries [a, b)
    if (a >= b) continue;
    // forced end on the last entry when its row goes on past the partition
    if (!is_end[(size_t)b - 1]) {
      is_end[(size_t)b - 1] = 2;
      target[(size_t)b - 1] = row_of[(size_t)b - 1];
    }
    // a row continuing from before: its first end here goes to the overflow row
    if (t > 0 && a > 0 && row_of[(size_t)a] == row_of[(size_t)a - 1]) {
      for (int64_t e = a; e < b; ++e)
        if (is_end[(size_t)e]) {
          target[(size_t)e] = m + t;
          break;
        }
      T.last_row = row_of[(size_t)a];
      T.overflow_row = (int)(m + t);
    }
  }
  // eor vectors, masks, row_arr
  auto& veceor = d.get_veceor_reference();
  auto& masks = d.get_scan_mask_reference();
  auto& rows = d.get_row_arr_reference();
  veceor.clear();
  masks.clear();
  rows.clear();
  for (int t = 0; t < nparts; ++t) {
    thr_info_t& T = thr[(size_t)t];
    T.vbase = (int)veceor.size();
    T.rbase = (int)rows.size();
    T.sbase = (int)(veceor.size() * (size_t)(nm + 1));
    for (int64_t v = T.start_vec; v < T.end_vec; ++v) {
      std::vector<uint8_t> endl((size_t)vec, 0);
      bool any = false;
      for (int j = 0; j < vec; ++j) {
        const int64_t e = v * vec + j;
        if (e < nnz && is_end[(size_t)e]) {
          endl[(size_t)j] = 1;
          any = true;
        }
      }
      if (!any) continue;
      veceor.push_back((uint32_t)v);
      std::vector<int> head((size_t)vec, 0);  // segment start of each lane
      for (int j = 1; j < vec; ++j) head[(size_t)j] = endl[(size_t)j - 1] ? j : head[(size_t)j - 1];
      for (int st = 0; st <= nm; ++st) {
        std::vector<uint8_t> mk((size_t)mb, 0);
        for (int j = 0; j < vec; ++j) {
          const bool bit = st < nm ? (j >= (1 << st) && j - (1 << st) >= head[(size_t)j])
                                   : endl[(size_t)j] != 0;
          if (bit) mk[(size_t)(j / 8)] |= (uint8_t)(1u << (j % 8));
        }
        masks.insert(masks.end(), mk.begin(), mk.end());
      }
      for (int j = 0; j < vec; ++j)
        if (endl[(size_t)j]) rows.push_back((uint32_t)target[(size_t)(v * vec + j)]);
    }
  }
  veceor.push_back((uint32_t)nvec);  // sentinel read after each partition's last eor vector
}

template void build_vhcc<double>(spMtxVHCC<double>&, const spMtxCRS<double>&, int, int);
template void build_vhcc<float>(spMtxVHCC<float>&, const spMtxCRS<float>&, int, int);

}  // namespace rvvla
