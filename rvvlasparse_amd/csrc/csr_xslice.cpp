// csr_xslice.cpp -- column-sliced device image of a CSR matrix.
//
// On gfx950 every XCD has its own L2 and workgroups are handed to the XCDs
// round-robin by workgroup index.  A matrix whose columns are spread over
// an x larger than one L2 (e.g. a power-law graph with uniform columns)
// turns every x gather into a miss to the memory-side cache.  The sliced
// image cuts the columns into P contiguous slices (P = the number of XCDs,
// or twice that) and re-lays the entries slice-major: slice p holds, for
// each row with entries in columns of slice p, those entries in CSR order.
// The multiply runs slice p only on the workgroups of XCD p mod X (X = the
// XCD count; workgroup b takes a block of slice b mod X of its phase), so
// each XCD's gathers stay inside one 1/P of x, which fits its L2.  With
// P = 2X the launch has two phases (slices 0..X-1, then X..2X-1) in
// workgroup order, which the hardware dispatches roughly in order.  Each
// (row, slice) pair leaves one partial sum; a second kernel adds a row's
// partials in slice order (deterministic) and applies alpha and beta.
//
// Arrays (pairs = number of (row, slice) pairs with at least one entry):
//   rst[pairs + 1]   entry offsets of the compressed rows, slice-major,
//                    rows ascending within a slice
//   col[nnz], val[nnz]
//   blk[...]         (first, end) compressed-row range of each workgroup:
//                    per phase h, nb_h * X entries, entry b of the phase in
//                    slice h*X + b mod X (empty ranges pad); within a slice
//                    the long-row blocks come first
//   mask[m]          (16 bits) bit p set when row r has entries in slice p
//   base[P + 1]      first pair of each slice
//   pref[P * nq]     present rows of slice p in rows [0, 64 q), nq = m/64
//                    rounded up; a wavefront of 64 rows finds the pair of
//                    row r in slice p as base[p] + pref[p*nq + q] + the
//                    number of lower lanes with bit p set.
#include <omp.h>

#include <algorithm>
#include <cmath>

#include "csr_blocks.h"
#include "internal.h"

namespace rvvla {

std::vector<int32_t> csr_row_blocks(const int* rst, int64_t r0, int64_t r1) {
  std::vector<int32_t> longs, streams;
  int64_t r = r0;
  while (r < r1) {
    if (rst[r + 1] - rst[r] > kLongRow) {
      longs.push_back((int32_t)r);
      ++r;
      continue;
    }
    const int64_t start = r;
    int64_t cnt = 0;
    while (r < r1 && r - start < kStreamRows) {
      const int64_t len = rst[r + 1] - rst[r];
      if (len > kLongRow || cnt + len > kStreamCap) break;
      cnt += len;
      ++r;
    }
    streams.push_back((int32_t)start);
    streams.push_back((int32_t)r);
  }
  // longest rows first: a long row is the longest-running block, so it
  // should not start last
  std::stable_sort(longs.begin(), longs.end(),
                   [&](int32_t a, int32_t b) { return rst[a + 1] - rst[a] > rst[b + 1] - rst[b]; });
  std::vector<int32_t> out;
  out.reserve(2 * longs.size() + streams.size());
  for (int32_t l : longs) {
    out.push_back(l);
    out.push_back(l + 1);
  }
  out.insert(out.end(), streams.begin(), streams.end());
  return out;
}

double far_fraction(int64_t m, int64_t n, const int* rst, const int* col, int64_t window) {
  if (m == 0 || rst[m] == 0) return 0.0;
  const double scale = (double)n / (double)m;
  int64_t far = 0;
#pragma omp parallel for schedule(static) reduction(+ : far)
  for (int64_t r = 0; r < m; ++r) {
    const double c0 = (double)r * scale;
    for (int e = rst[r]; e < rst[r + 1]; ++e)
      far += std::abs((double)col[e] - c0) > (double)window;
  }
  return (double)far / (double)rst[m];
}

template <class V>
void build_xslice(XSliceHost<V>& X, int64_t m, int64_t n, const int* rst, const int* col,
                  const V* val, int P, int xcd) {
  if (P < 2 || P > 16) throw Error(RVVLA_ERR_INVALID_ARG, "column slices: P in [2, 16]");
  if (xcd < 1 || xcd > P) xcd = P;
  const int64_t nnz = m ? rst[m] : 0;
  const int64_t w = (n + P - 1) / P;
  X.P = P;
  X.width = w;
  X.mask.assign((size_t)m, 0);
  // per row: entries in each slice
  std::vector<int32_t> cnt((size_t)(m * P), 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t r = 0; r < m; ++r) {
    int32_t* c = &cnt[(size_t)(r * P)];
    for (int e = rst[r]; e < rst[r + 1]; ++e) c[col[e] / w]++;
    uint16_t mk = 0;
    for (int p = 0; p < P; ++p) mk |= (uint16_t)((c[p] > 0) << p);
    X.mask[(size_t)r] = mk;
  }
  // pairs per slice and per 64-row group
  const int64_t nq = (m + 63) / 64;
  X.nq = nq;
  X.pref.assign((size_t)(P * nq), 0);
  X.base.assign((size_t)P + 1, 0);
  for (int p = 0; p < P; ++p) {
    int64_t acc = 0;
    for (int64_t q = 0; q < nq; ++q) {
      X.pref[(size_t)(p * nq + q)] = (int32_t)acc;
      const int64_t hi = std::min(m, (q + 1) * 64);
      for (int64_t r = q * 64; r < hi; ++r) acc += (X.mask[(size_t)r] >> p) & 1;
    }
    X.base[(size_t)p + 1] = X.base[(size_t)p] + (int32_t)acc;
  }
  const int64_t pairs = X.base[(size_t)P];
  // compressed row pointers, slice-major
  X.rst.assign((size_t)pairs + 1, 0);
  {
    int64_t k = 0;
    for (int p = 0; p < P; ++p)
      for (int64_t r = 0; r < m; ++r) {
        const int32_t c = cnt[(size_t)(r * P + p)];
        if (c) {
          X.rst[(size_t)k + 1] = X.rst[(size_t)k] + c;
          ++k;
        }
      }
  }
  // entries: the pair of row r in slice p is base[p] + rank of r among the
  // rows present in slice p
  X.col.assign((size_t)nnz, 0);
  X.val.assign((size_t)nnz, V(0));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < nq; ++q) {
    int32_t rank[16];
    for (int p = 0; p < P; ++p) rank[p] = X.base[(size_t)p] + X.pref[(size_t)(p * nq + q)];
    const int64_t hi = std::min(m, (q + 1) * 64);
    for (int64_t r = q * 64; r < hi; ++r) {
      int32_t fill[16];
      for (int p = 0; p < P; ++p) fill[p] = X.rst[(size_t)rank[p]];
      for (int e = rst[r]; e < rst[r + 1]; ++e) {
        const int p = (int)(col[e] / w);
        X.col[(size_t)fill[p]] = col[e];
        X.val[(size_t)fill[p]++] = val[e];
      }
      for (int p = 0; p < P; ++p) rank[p] += (X.mask[(size_t)r] >> p) & 1;
    }
  }
  // workgroup blocks: per slice, then interleaved phase by phase so that
  // entry b of a phase belongs to slice (phase * xcd + b mod xcd)
  std::vector<std::vector<int32_t>> per((size_t)P);
  for (int p = 0; p < P; ++p)
    per[(size_t)p] = csr_row_blocks(X.rst.data(), X.base[(size_t)p], X.base[(size_t)p + 1]);
  X.blk.clear();
  for (int p0 = 0; p0 < P; p0 += xcd) {
    const int np = std::min(xcd, P - p0);
    size_t nb = 0;
    for (int p = p0; p < p0 + np; ++p) nb = std::max(nb, per[(size_t)p].size() / 2);
    for (size_t i = 0; i < nb; ++i)
      for (int j = 0; j < xcd; ++j) {
        const int p = p0 + j;
        int32_t lo, hi;
        if (j < np && 2 * i + 1 < per[(size_t)p].size()) {
          lo = per[(size_t)p][2 * i];
          hi = per[(size_t)p][2 * i + 1];
        } else {
          lo = hi = X.base[(size_t)std::min(p, P)];  // empty
        }
        X.blk.push_back(lo);
        X.blk.push_back(hi);
      }
  }
}

template void build_xslice<double>(XSliceHost<double>&, int64_t, int64_t, const int*, const int*,
                                   const double*, int, int);
template void build_xslice<float>(XSliceHost<float>&, int64_t, int64_t, const int*, const int*,
                                  const float*, int, int);

}  // namespace rvvla
