// convert_sell.cpp -- CRS -> SELL-C-sigma host builder.
//
// The reference ships no SELL builder; the layout here is the one its
// kernels read (reference src/scalar/spmv_mtx_sell_c_sigma.cpp:76-95 and
// src/rvv/spmv_mtx_sell_c_sigma.cpp:60-85): slice i covers sorted rows
// SBs[i] .. SBs[i]+Bs[i]-1, has width Cl[i] (its longest row) and stores
// entry (slot j, row k) at Cs[i] + j*Bs[i] + k.  The sort follows Kreutzer
// et al. (SIAM J. Sci. Comput. 2014): rows are stable-sorted by descending
// length inside consecutive windows of sigma rows.  Padding slots carry
// value 0 and the row's last real column (column 0 for an empty row), so
// every slot is a valid gather.
#include <omp.h>

#include <algorithm>
#include <numeric>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

int default_sell_sigma(int64_t m, int64_t nnz, const int* rst) {
  // sigma = 1 keeps the row order (the permutation is the identity, so the
  // multiply needs no vPerm read): taken when 64-row slices in natural
  // order pad by at most 1 % of the entries.  Otherwise sorting all rows by
  // length (sigma = m) when that at least halves the padded slots (matrices
  // with very uneven rows), else windows of 16384 rows, which keep the y
  // writes of a slice close together.
  if (m == 0 || nnz == 0 || !rst) return 16384;
  int64_t natural = 0;
  for (int64_t r0 = 0; r0 < m; r0 += 64) {
    const int64_t r1 = std::min(m, r0 + 64);
    int64_t w = 0;
    for (int64_t r = r0; r < r1; ++r) w = std::max<int64_t>(w, rst[r + 1] - rst[r]);
    natural += w * (r1 - r0);
  }
  if ((double)(natural - nnz) <= 0.01 * (double)nnz) return 1;
  return (double)sorted_sell_slots(m, rst) <= 0.5 * (double)natural ? (int)m : 16384;
}

int64_t sorted_sell_slots(int64_t m, const int* rst) {
  // slice k of the rows sorted by descending length starts with its widest
  std::vector<int32_t> len((size_t)m);
  for (int64_t r = 0; r < m; ++r) len[(size_t)r] = rst[r + 1] - rst[r];
  std::sort(len.begin(), len.end(), [](int32_t a, int32_t b) { return a > b; });
  int64_t slots = 0;
  for (int64_t r0 = 0; r0 < m; r0 += 64) slots += (int64_t)len[(size_t)r0] * std::min<int64_t>(64, m - r0);
  return slots;
}

template <class V>
void build_sell(spMtxSELL_C_Sigma<V>& d, const spMtxCRS<V>& s, int C, int sigma) {
  if (C <= 0) throw Error(RVVLA_ERR_INVALID_ARG, "SELL: C must be positive");
  if (sigma <= 0) throw Error(RVVLA_ERR_INVALID_ARG, "SELL: sigma must be positive");
  const int64_t m = (int64_t)s.m;
  d.freeMem();
  d.m = s.m;
  d.n = s.n;
  d.nz = s.nz;
  d.C = C;
  d.Sigma = sigma;
  d.cnt_b = (size_t)((m + C - 1) / C);
  d.vPerm.resize((size_t)m);
  std::iota(d.vPerm.begin(), d.vPerm.end(), 0u);
  auto len = [&](uint32_t r) { return s.Rst[r + 1] - s.Rst[r]; };
  if (sigma > 1) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w0 = 0; w0 < m; w0 += sigma) {
      const int64_t w1 = std::min<int64_t>(m, w0 + sigma);
      std::stable_sort(d.vPerm.begin() + w0, d.vPerm.begin() + w1,
                       [&](uint32_t a, uint32_t b) { return len(a) > len(b); });
    }
  }
  const size_t nb = d.cnt_b;
  d.vCl.assign(nb, 0);
  d.vCs.assign(nb, 0);
  d.vBs.assign(nb, 0);
  d.vSBs.assign(nb, 0);
  for (size_t i = 0; i < nb; ++i) {
    const int64_t r0 = (int64_t)i * C, r1 = std::min<int64_t>(m, r0 + C);
    int w = 0;
    for (int64_t k = r0; k < r1; ++k) w = std::max(w, len(d.vPerm[(size_t)k]));
    d.vCl[i] = w;
    d.vBs[i] = (int)(r1 - r0);
    d.vSBs[i] = (int)r0;
  }
  int64_t slots = 0;
  for (size_t i = 0; i < nb; ++i) {
    if (slots > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "SELL: more than 2^31 slots");
    d.vCs[i] = (int)slots;
    slots += (int64_t)d.vCl[i] * d.vBs[i];
  }
  if (slots > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "SELL: more than 2^31 slots");
  d.vCol.assign((size_t)slots, 0);
  d.vVal.assign((size_t)slots, V(0));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)nb; ++i) {
    const int bs = d.vBs[(size_t)i], w = d.vCl[(size_t)i];
    const int64_t cs = d.vCs[(size_t)i];
    for (int k = 0; k < bs; ++k) {
      const uint32_t r = d.vPerm[(size_t)(d.vSBs[(size_t)i] + k)];
      const int e0 = s.Rst[r], l = s.Rst[r + 1] - e0;
      const uint32_t pad_col = l > 0 ? (uint32_t)s.Col[e0 + l - 1] : 0u;
      for (int j = 0; j < w; ++j) {
        const int64_t slot = cs + (int64_t)j * bs + k;
        if (j < l) {
          d.vCol[(size_t)slot] = (uint32_t)s.Col[e0 + j];
          d.vVal[(size_t)slot] = s.Val[e0 + j];
        } else {
          d.vCol[(size_t)slot] = pad_col;
        }
      }
    }
  }
}

template void build_sell<double>(spMtxSELL_C_Sigma<double>&, const spMtxCRS<double>&, int, int);
template void build_sell<float>(spMtxSELL_C_Sigma<float>&, const spMtxCRS<float>&, int, int);

}  // namespace rvvla
