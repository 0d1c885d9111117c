// convert_lav.cpp -- CRS -> LAV host builder.
//
// The reference ships no LAV builder; the layout is the one its kernels
// read (reference src/scalar/spmv_mtx_lav.cpp:52-79, include/LAV.h:13-31),
// built after Gomez et al. ("Efficiently running SpMV on long vector
// architectures", PPoPP 2021):
//  * columns are ranked by entry count (descending, ties by index); the
//    smallest prefix of ranked columns holding at least (1 - T) of the
//    entries is the dense part, cut by rank into segments of segment_len
//    columns; all other entries form sparse_part (CSR, entry order kept);
//  * rows are stable-sorted once by their dense-part length (descending) and
//    cut into chunk_count chunks of SIMD_Lanes rows; the same row order is
//    used in every segment (out_order repeats it per segment);
//  * in segment s, chunk c takes as many lane steps as its longest row has
//    entries there; a row's entries in a segment keep their CSR order; steps
//    below the chunk's shortest row use all lanes, later steps record in
//    mask[] the number of leading lanes that may hold entries (1 + the
//    highest lane with an entry); slots without an entry hold value 0 and
//    the row's last column (column 0 for a row with none).
#include <omp.h>

#include <algorithm>
#include <numeric>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

template <class V>
void build_lav(spMtxLAV<V>& d, const spMtxCRS<V>& s, int lanes, double T, int segment_len) {
  if (lanes < 1 || lanes > 64) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: SIMD_Lanes in [1, 64]");
  if (segment_len < 1) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: segment_len must be positive");
  if (!(T >= 0.0 && T <= 1.0)) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: T in [0, 1]");
  const int64_t m = (int64_t)s.m, n = (int64_t)s.n, nnz = (int64_t)s.nz, L = lanes;
  d = spMtxLAV<V>();
  d.SIMD_Lanes = (size_t)L;
  d.m = (int)m;
  d.n = (int)n;
  d.nz = (int)nnz;
  d.T = T;
  d.segment_len = segment_len;
  // column ranking
  std::vector<int64_t> cnt((size_t)n, 0);
  for (int64_t e = 0; e < nnz; ++e) cnt[(size_t)s.Col[e]]++;
  std::vector<int32_t> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return cnt[(size_t)a] > cnt[(size_t)b]; });
  int64_t ndense = 0, acc = 0;
  const double need = (1.0 - T) * (double)nnz;
  while (ndense < n && (double)acc < need && cnt[(size_t)order[(size_t)ndense]] > 0)
    acc += cnt[(size_t)order[(size_t)ndense++]];
  std::vector<int32_t> seg_of((size_t)n, -1);
  for (int64_t k = 0; k < ndense; ++k) seg_of[(size_t)order[(size_t)k]] = (int32_t)(k / segment_len);
  const int64_t S = (ndense + segment_len - 1) / segment_len;
  d.segment_count = (size_t)S;
  const int64_t cc = (m + L - 1) / L;
  d.chunk_count = (size_t)cc;
  // sparse part
  {
    std::vector<int> rst((size_t)m + 1, 0);
    for (int64_t r = 0; r < m; ++r) {
      int c = 0;
      for (int e = s.Rst[r]; e < s.Rst[r + 1]; ++e) c += seg_of[(size_t)s.Col[e]] < 0;
      rst[(size_t)r + 1] = rst[(size_t)r] + c;
    }
    spMtxCRS<V> sp((size_t)m, (size_t)n, (size_t)rst[(size_t)m]);
    std::copy(rst.begin(), rst.end(), sp.Rst);
    int64_t k = 0;
    for (int64_t e = 0; e < nnz; ++e)
      if (seg_of[(size_t)s.Col[e]] < 0) {
        sp.Col[k] = s.Col[e];
        sp.Val[k++] = s.Val[e];
      }
    d.sparse_part = std::move(sp);
  }
  // per row: dense entries grouped by segment (CSR order inside a segment)
  std::vector<int64_t> rcnt((size_t)(m * std::max<int64_t>(S, 1)), 0);  // [r*S + s]
  std::vector<int64_t> dlen((size_t)m, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; ++r)
    for (int e = s.Rst[r]; e < s.Rst[r + 1]; ++e) {
      const int32_t g = seg_of[(size_t)s.Col[e]];
      if (g >= 0) {
        rcnt[(size_t)(r * S + g)]++;
        dlen[(size_t)r]++;
      }
    }
  std::vector<uint32_t> perm((size_t)m);
  std::iota(perm.begin(), perm.end(), 0u);
  std::stable_sort(perm.begin(), perm.end(),
                   [&](uint32_t a, uint32_t b) { return dlen[a] > dlen[b]; });
  const uint32_t pad_row = m > 0 ? perm[(size_t)m - 1] : 0u;
  d.out_order.assign((size_t)(S * cc * L), 0u);
  for (int64_t g = 0; g < S; ++g)
    for (int64_t i = 0; i < cc * L; ++i)
      d.out_order[(size_t)((g * cc) * L + i)] = i < m ? perm[(size_t)i] : pad_row;
  // steps per (segment, chunk)
  std::vector<int64_t> steps((size_t)(S * cc)), full((size_t)(S * cc));
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < S * cc; ++q) {
    const int64_t g = q / cc, c = q % cc;
    int64_t mx = 0, mn = INT64_MAX;
    for (int64_t i = 0; i < L; ++i) {
      const int64_t row_pos = c * L + i;
      const int64_t k = row_pos < m ? rcnt[(size_t)(perm[(size_t)row_pos] * S + g)] : 0;
      mx = std::max(mx, k);
      mn = std::min(mn, k);
    }
    steps[(size_t)q] = mx;
    full[(size_t)q] = mn;
  }
  d.chunk_offsets.assign((size_t)(S * (cc + 1)), 0);
  d.chunk_offsets_full_mask.assign((size_t)(S * cc), 0);
  d.segment_ptr.assign((size_t)S + 1, 0);
  int64_t total = 0;
  for (int64_t g = 0; g < S; ++g) {
    d.segment_ptr[(size_t)g] = (size_t)total;
    int64_t off = 0;
    for (int64_t c = 0; c < cc; ++c) {
      d.chunk_offsets[(size_t)(g * (cc + 1) + c)] = (int)off;
      d.chunk_offsets_full_mask[(size_t)(g * cc + c)] = (int)(off + full[(size_t)(g * cc + c)]);
      off += steps[(size_t)(g * cc + c)];
    }
    d.chunk_offsets[(size_t)(g * (cc + 1) + cc)] = (int)off;
    total += off;
    if (total * L > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "LAV: dense part too large");
  }
  d.segment_ptr[(size_t)S] = (size_t)total;
  d.mask.assign((size_t)total, (uint16_t)L);
  d.val.assign((size_t)(total * L), V(0));
  d.col_id.assign((size_t)(total * L), 0u);
  // fill: row r's k-th entry of segment g goes to step off + k, lane i
#pragma omp parallel
  {
    std::vector<int64_t> pos((size_t)std::max<int64_t>(S, 1));
#pragma omp for schedule(dynamic, 16)
    for (int64_t c = 0; c < cc; ++c) {
      for (int64_t i = 0; i < L; ++i) {
        const int64_t row_pos = c * L + i;
        if (row_pos >= m) continue;
        const uint32_t r = perm[(size_t)row_pos];
        for (int64_t g = 0; g < S; ++g)
          pos[(size_t)g] = (int64_t)d.segment_ptr[(size_t)g] + d.chunk_offsets[(size_t)(g * (cc + 1) + c)];
        uint32_t last_col = 0;
        for (int e = s.Rst[r]; e < s.Rst[r + 1]; ++e) {
          const int32_t g = seg_of[(size_t)s.Col[e]];
          if (g < 0) continue;
          const int64_t slot = pos[(size_t)g]++ * L + i;
          d.val[(size_t)slot] = s.Val[e];
          d.col_id[(size_t)slot] = (uint32_t)s.Col[e];
          last_col = (uint32_t)s.Col[e];
        }
        // pads of this lane: the row's last dense column
        for (int64_t g = 0; g < S; ++g) {
          const int64_t a = (int64_t)d.segment_ptr[(size_t)g] + d.chunk_offsets[(size_t)(g * (cc + 1) + c)];
          const int64_t b = (int64_t)d.segment_ptr[(size_t)g] + d.chunk_offsets[(size_t)(g * (cc + 1) + c + 1)];
          for (int64_t k = pos[(size_t)g]; k < b; ++k) d.col_id[(size_t)(k * L + i)] = last_col;
          (void)a;
        }
      }
      // masks of the partial steps
      for (int64_t g = 0; g < S; ++g) {
        const int64_t base = (int64_t)d.segment_ptr[(size_t)g];
        const int64_t a = d.chunk_offsets_full_mask[(size_t)(g * cc + c)];
        const int64_t b = d.chunk_offsets[(size_t)(g * (cc + 1) + c + 1)];
        for (int64_t k = a; k < b; ++k) {
          int hi = 0;
          for (int64_t i = 0; i < L; ++i) {
            const int64_t row_pos = c * L + i;
            if (row_pos < m && rcnt[(size_t)(perm[(size_t)row_pos] * S + g)] >
                                   k - d.chunk_offsets[(size_t)(g * (cc + 1) + c)])
              hi = (int)i + 1;
          }
          d.mask[(size_t)(base + k)] = (uint16_t)hi;
        }
      }
    }
  }
}

template void build_lav<double>(spMtxLAV<double>&, const spMtxCRS<double>&, int, double, int);
template void build_lav<float>(spMtxLAV<float>&, const spMtxCRS<float>&, int, double, int);

}  // namespace rvvla
