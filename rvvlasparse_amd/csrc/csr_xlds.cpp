// csr_xlds.cpp -- LDS-sliced device image of a CSR matrix.
//
// For a matrix whose columns are spread over an x far larger than the L2
// (a power-law graph with uniform columns), each x gather costs a whole
// cache line from L2 or the memory-side cache.  This image cuts x into
// slices of W = 16384 elements that fit one workgroup's LDS (128 KiB of
// doubles, 64 KiB of floats) and re-lays the entries slice-major: slice p
// holds, for each row with entries in columns [p W, (p + 1) W), those
// entries in CSR order, as 16-bit column offsets inside the slice.  Each
// workgroup (as many per CU as its LDS allows) takes an equal share of the
// slice-major entries; it
// loads each slice of x it meets into LDS once (a coalesced read) and
// streams the entries, gathering x from LDS.  Each (row, slice) pair leaves
// one partial sum.  The partials are grouped by row blocks (at most
// comb_rows rows and comb_pairs pairs, or one longer row); inside a
// block they are ordered by slice, then row, so that each slice writes
// runs of consecutive slots.  A second kernel loads a block's partials,
// puts each at its row-major place in LDS and adds each row's partials in
// slice order (deterministic), then applies alpha and beta.
//
// Arrays (pairs = number of (row, slice) pairs with at least one entry,
// numbered slice-major, rows ascending within a slice):
//   col[nnz] (uint16), val[nnz]   entries, slice-major; bits 14..0 the
//                                 column inside the slice (W <= 32768),
//                                 bit 15 set on the first entry of a pair
//   pst[pairs + 1]                entry offset of each pair
//   wpos[pairs]                   slot of each pair's partial
//   rowptr[m + 1]                 slots of row r's block-row-major places:
//                                 rowptr[r] .. rowptr[r+1]-1, slices
//                                 ascending
//   lpos[pairs] (uint16)          for slot q of a block starting at slot b:
//                                 the row-major place (minus b) of the
//                                 partial stored there
//   cblk[4 * nblocks]             (first row, end row, first slot, slots)
//                                 of each combine block
//   units[3 * nunits]             (slice, first pair, end pair), contiguous
//                                 per workgroup; wg[w] .. wg[w+1]-1 are the
//                                 units of workgroup w
#include <omp.h>

#include <algorithm>

#include "csr_blocks.h"
#include "internal.h"

namespace rvvla {

template <class V>
void build_xlds(XLdsHost<V>& X, int64_t m, int64_t n, const int* rst, const int* col, const V* val,
                int64_t W, int64_t nwg, int64_t comb_rows, int64_t comb_pairs) {
  const int64_t nnz = m > 0 ? rst[m] : 0;
  const int64_t S = n > 0 ? (n + W - 1) / W : 0;
  nwg = std::max<int64_t>(1, nwg);
  X.W = W;
  X.S = S;
  X.rowptr.assign((size_t)m + 1, 0);
  // pairs of each row (columns ascending within a row: one pair per run)
  std::vector<int32_t> rp((size_t)m, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; ++r) {
    int32_t c = 0;
    int64_t last = -1;
    for (int e = rst[r]; e < rst[r + 1]; ++e) {
      const int64_t p = col[e] / W;
      if (p != last) ++c;
      last = p;
    }
    rp[(size_t)r] = c;
  }
  for (int64_t r = 0; r < m; ++r) X.rowptr[(size_t)r + 1] = X.rowptr[(size_t)r] + rp[(size_t)r];
  const int64_t npairs = X.rowptr[(size_t)m];
  X.npairs = npairs;
  // per (row chunk, slice) entry and pair counts -> slice-major offsets
  const int T = std::max(1, std::min<int>(omp_get_max_threads(), (int)std::max<int64_t>(m / 4096, 1)));
  std::vector<int64_t> ce((size_t)T * S, 0), cp((size_t)T * S, 0);
  auto chunk = [&](int t) { return std::make_pair(m * t / T, m * (t + 1) / T); };
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    const auto rr = chunk(t);
    int64_t* e = ce.data() + (size_t)t * S;
    int64_t* q = cp.data() + (size_t)t * S;
    for (int64_t r = rr.first; r < rr.second; ++r) {
      int64_t last = -1;
      for (int k = rst[r]; k < rst[r + 1]; ++k) {
        const int64_t p = col[k] / W;
        e[p]++;
        if (p != last) q[p]++;
        last = p;
      }
    }
  }
  std::vector<int64_t> pbase((size_t)S + 1, 0);
  {
    int64_t eo = 0, po = 0;
    for (int64_t p = 0; p < S; ++p) {
      pbase[(size_t)p] = po;
      for (int t = 0; t < T; ++t) {
        const int64_t ne = ce[(size_t)t * S + p], np = cp[(size_t)t * S + p];
        ce[(size_t)t * S + p] = eo;
        cp[(size_t)t * S + p] = po;
        eo += ne;
        po += np;
      }
    }
    pbase[(size_t)S] = po;
  }
  X.col.assign((size_t)nnz, 0);
  X.val.assign((size_t)nnz, V(0));
  X.pst.assign((size_t)npairs + 1, 0);
  std::vector<int32_t> dest((size_t)npairs, 0), slot_slice((size_t)npairs, 0);
  X.pst[(size_t)npairs] = (int32_t)nnz;
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    const auto rr = chunk(t);
    int64_t* e = ce.data() + (size_t)t * S;
    int64_t* q = cp.data() + (size_t)t * S;
    for (int64_t r = rr.first; r < rr.second; ++r) {
      int64_t last = -1, k = 0;
      for (int i = rst[r]; i < rst[r + 1]; ++i) {
        const int64_t p = col[i] / W;
        const bool first = p != last;
        if (first) {
          const int64_t g = q[p]++;
          const int32_t slot = X.rowptr[(size_t)r] + (int32_t)k++;
          X.pst[(size_t)g] = (int32_t)e[p];
          dest[(size_t)g] = slot;
          slot_slice[(size_t)slot] = (int32_t)p;
          last = p;
        }
        const int64_t at = e[p]++;
        X.col[(size_t)at] = (uint16_t)((col[i] - p * W) | (first ? 0x8000 : 0));
        X.val[(size_t)at] = val[i];
      }
    }
  }
  // combine blocks and the block-local (slice, row) order of the slots
  X.cblk.clear();
  for (int64_t r = 0; r < m;) {
    const int64_t r0 = r;
    if (rp[(size_t)r] > comb_pairs) {
      ++r;
    } else {
      while (r < m && r - r0 < comb_rows && X.rowptr[(size_t)r + 1] - X.rowptr[(size_t)r0] <= comb_pairs)
        ++r;
    }
    X.cblk.push_back((int32_t)r0);
    X.cblk.push_back((int32_t)r);
    X.cblk.push_back(X.rowptr[(size_t)r0]);
    X.cblk.push_back(X.rowptr[(size_t)r] - X.rowptr[(size_t)r0]);
  }
  const int64_t nb = (int64_t)X.cblk.size() / 4;
  std::vector<int32_t> wslot((size_t)npairs, 0);
  X.lpos.assign((size_t)npairs, 0);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int32_t>> ord;
#pragma omp for schedule(dynamic, 64)
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t base = X.cblk[(size_t)(4 * b + 2)];
      const int64_t end = base + X.cblk[(size_t)(4 * b + 3)];
      ord.clear();
      for (int64_t q = base; q < end; ++q) ord.emplace_back(slot_slice[(size_t)q], (int32_t)(q - base));
      std::stable_sort(ord.begin(), ord.end(),
                       [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& c) {
                         return a.first < c.first;
                       });
      for (size_t i = 0; i < ord.size(); ++i) {
        wslot[(size_t)(base + ord[i].second)] = (int32_t)(base + (int64_t)i);
        X.lpos[(size_t)(base + (int64_t)i)] = (uint16_t)(ord[i].second & 0xFFFF);
      }
    }
  }
  X.wpos.resize((size_t)npairs);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < npairs; ++g) X.wpos[(size_t)g] = wslot[(size_t)dest[(size_t)g]];
  // units: the slice-major pairs cut into nwg runs of about nnz / nwg
  // entries at pair boundaries, each run cut again at slice boundaries
  X.units.clear();
  X.wg.assign((size_t)nwg + 1, 0);
  const int64_t target = std::max<int64_t>(1, (nnz + nwg - 1) / nwg);
  int64_t w = 0, acc = 0;
  for (int64_t p = 0; p < S; ++p) {
    int64_t g = pbase[(size_t)p];
    const int64_t g1 = pbase[(size_t)p + 1];
    while (g < g1) {
      if (acc >= target && w + 1 < nwg) {
        X.wg[(size_t)++w] = (int32_t)(X.units.size() / 3);
        acc = 0;
      }
      const int64_t start = g;
      while (g < g1 && (g == start || acc < target || w + 1 == nwg)) {
        acc += X.pst[(size_t)g + 1] - X.pst[(size_t)g];
        ++g;
      }
      X.units.push_back((int32_t)p);
      X.units.push_back((int32_t)start);
      X.units.push_back((int32_t)g);
    }
  }
  for (int64_t k = w + 1; k <= nwg; ++k) X.wg[(size_t)k] = (int32_t)(X.units.size() / 3);
}

int64_t xlds_pairs(int64_t m, const int* rst, const int* col, int64_t W) {
  int64_t pairs = 0;
#pragma omp parallel for schedule(static) reduction(+ : pairs)
  for (int64_t r = 0; r < m; ++r) {
    int64_t last = -1;
    for (int e = rst[r]; e < rst[r + 1]; ++e) {
      const int64_t p = col[e] / W;
      pairs += p != last;
      last = p;
    }
  }
  return pairs;
}

template void build_xlds<double>(XLdsHost<double>&, int64_t, int64_t, const int*, const int*,
                                 const double*, int64_t, int64_t, int64_t, int64_t);
template void build_xlds<float>(XLdsHost<float>&, int64_t, int64_t, const int*, const int*,
                                const float*, int64_t, int64_t, int64_t, int64_t);

}  // namespace rvvla
