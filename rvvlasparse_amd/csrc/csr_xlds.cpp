// csr_xlds.cpp -- LDS-sliced device image of a CSR matrix.
//
// For a matrix whose columns are spread over an x far larger than the L2
// (a power-law graph with uniform columns), each x gather costs a whole
// cache line from L2 or the memory-side cache.  This image cuts x into
// slices of W elements that fill one workgroup's LDS (128 KiB: W = 16384
// doubles or 32768 floats) and re-lays the entries slice-major: slice p
// holds, for each row with entries in columns [p W, (p + 1) W), those
// entries in CSR order, as 16-bit column offsets inside the slice.
//
// The slice-major entries of each slice are cut into tasks of kXlTask
// entries (the last task of a slice may be shorter); each task takes
// kXlTask slots.  One wavefront runs one task: lane l takes the 8
// consecutive entries l*8 .. l*8+7 of it.  Inside a task the values are
// stored so that the wavefront's 16-byte loads are contiguous: with vpw
// values to 16 bytes, entry 8 l + k of the task sits at slot
// ((k / vpw) * 64 + l) * vpw + k % vpw (the identity for 16-bit values).
// A (row, slice) pair's entries are cut at task boundaries into pieces;
// each piece leaves one partial sum.  Bit 15
// of an entry's column marks the first entry of a piece, so every task
// starts with a marked entry and no partial sum crosses a task.  Pieces
// are numbered in stream order (slice-major, rows ascending inside a
// slice), which is also the order of their partial slots: the slice pass
// writes the partials of a task to consecutive slots.
//
// The combine pass works on row blocks (at most comb_rows rows and
// comb_pieces pieces, or one row with more pieces): for block b and slice
// p the block's pieces are the slots off[b S + p] .. off[(b + 1) S + p] - 1
// (rows ascending inside a slice), so a block reads one run of slots per
// slice.  lpos gives each slot its row-major place inside the block (a
// row's pieces in stream order, i.e. slices ascending), where the combine
// puts it in LDS before adding each row's partials in that order.
//
// Arrays (pieces = partial slots):
//   col[nent] (uint16), val[nent]  entries slice-major, kXlTask slots a
//                                  task (values permuted inside a task as
//                                  above); bits 14..0 of col the column
//                                  inside the slice, bit 15 set on the
//                                  first entry of a piece; padding slots
//                                  are zero
//   task[4 * ntasks]               (first entry, entries, first piece,
//                                  slice) of each task
//   units[3 * nunits]              (slice, first task, end task),
//                                  contiguous per workgroup; wg[w] ..
//                                  wg[w+1]-1 are the units of workgroup w
//   rowptr[m + 1]                  pieces of the rows (prefix)
//   cblk[2 * nblocks]              (first row, end row) of each block
//   off[(nblocks + 1) * S]         first slot of slice p in block b
//   lpos[pieces] (uint16)          block-local row-major place of a slot
#include <omp.h>

#include <algorithm>
#include <stdexcept>

#include "internal.h"

namespace rvvla {

namespace {
inline int64_t round_task(int64_t v) { return (v + kXlTask - 1) / kXlTask * kXlTask; }
}  // namespace

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

int64_t xl_value_slot(int64_t at, int vpw) {
  const int64_t r = at % kXlTask, lane = r / 8, k = r % 8;
  return at - r + ((k / vpw) * 64 + lane) * vpw + k % vpw;
}

template <class V>
void build_xlds(XLdsHost<V>& X, int64_t m, int64_t n, const int* rst, const int* col, const V* val,
                int64_t W, int64_t nwg, int64_t comb_rows, int64_t comb_pieces, int vpw) {
  constexpr int64_t TE = kXlTask;
  const int64_t nnz = m > 0 ? rst[m] : 0;
  const int64_t S = n > 0 ? (n + W - 1) / W : 0;
  nwg = std::max<int64_t>(1, nwg);
  X = XLdsHost<V>();
  X.W = W;
  X.S = S;
  // ---- pairs of each row, and entries / pairs per (row chunk, slice)
  std::vector<int32_t> rpairs((size_t)m, 0);
  const int T = std::max(1, std::min<int>(omp_get_max_threads(), (int)std::max<int64_t>(m / 4096, 1)));
  auto chunk = [&](int t) { return std::make_pair(m * t / T, m * (t + 1) / T); };
  std::vector<int64_t> ce((size_t)T * S, 0), cp((size_t)T * S, 0);
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    const auto rr = chunk(t);
    int64_t* e = ce.data() + (size_t)t * S;
    int64_t* q = cp.data() + (size_t)t * S;
    for (int64_t r = rr.first; r < rr.second; ++r) {
      int64_t last = -1;
      int32_t c = 0;
      for (int k = rst[r]; k < rst[r + 1]; ++k) {
        const int64_t p = col[k] / W;
        e[p]++;
        if (p != last) {
          q[p]++;
          ++c;
        }
        last = p;
      }
      rpairs[(size_t)r] = c;
    }
  }
  // slice segments (rounded up to whole tasks), per-chunk offsets
  std::vector<int64_t> sb((size_t)S + 1, 0), pb((size_t)S + 1, 0);
  for (int64_t p = 0; p < S; ++p) {
    int64_t ne = 0, np = 0;
    for (int t = 0; t < T; ++t) {
      const int64_t a = ce[(size_t)t * S + p], b = cp[(size_t)t * S + p];
      ce[(size_t)t * S + p] = sb[(size_t)p] + ne;
      cp[(size_t)t * S + p] = pb[(size_t)p] + np;
      ne += a;
      np += b;
    }
    sb[(size_t)p + 1] = sb[(size_t)p] + round_task(ne);
    pb[(size_t)p + 1] = pb[(size_t)p] + np;
  }
  std::vector<int64_t> sne((size_t)S, 0);  // true entries of each slice (after the fill)
  X.nent = std::max<int64_t>(S > 0 ? sb[(size_t)S] : 0, 1);
  if (X.nent > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "CSR: LDS-sliced image too large");
  std::vector<int64_t> rpp((size_t)m + 1, 0);  // row-major pair pointer
  for (int64_t r = 0; r < m; ++r) rpp[(size_t)r + 1] = rpp[(size_t)r] + rpairs[(size_t)r];
  const int64_t npairs = m > 0 ? rpp[(size_t)m] : 0;
  // ---- entries slice-major; per row-major pair: position, slice-major
  // index, slice and length
  X.col.assign((size_t)X.nent, 0);
  X.val.assign((size_t)X.nent, V(0));
  std::vector<int32_t> ppos((size_t)npairs), pidx((size_t)npairs), pslc((size_t)npairs),
      plen((size_t)npairs);
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    const auto rr = chunk(t);
    int64_t* e = ce.data() + (size_t)t * S;
    int64_t* q = cp.data() + (size_t)t * S;
    for (int64_t r = rr.first; r < rr.second; ++r) {
      int64_t last = -1, k = rpp[(size_t)r] - 1;
      for (int i = rst[r]; i < rst[r + 1]; ++i) {
        const int64_t p = col[i] / W;
        const int64_t at = e[p]++;
        const bool first = p != last;
        if (first) {
          ++k;
          ppos[(size_t)k] = (int32_t)at;
          pidx[(size_t)k] = (int32_t)q[p]++;
          pslc[(size_t)k] = (int32_t)p;
          plen[(size_t)k] = 0;
          last = p;
        }
        plen[(size_t)k]++;
        X.col[(size_t)at] = (uint16_t)((col[i] - p * W) | (first ? 0x8000 : 0));
        X.val[(size_t)xl_value_slot(at, vpw)] = val[i];
      }
    }
  }
  // true entry count of each slice: the last chunk's running offset
  for (int64_t p = 0; p < S; ++p) sne[(size_t)p] = ce[(size_t)(T - 1) * S + p] - sb[(size_t)p];
  // ---- tasks: each slice's entries cut every TE; a task start inside a
  // pair starts a new piece of that pair
  std::vector<int64_t> tf((size_t)S + 1, 0);  // first task of each slice
  for (int64_t p = 0; p < S; ++p) tf[(size_t)p + 1] = tf[(size_t)p] + (sne[(size_t)p] + TE - 1) / TE;
  const int64_t ntasks = S > 0 ? tf[(size_t)S] : 0;
  if (ntasks > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "CSR: LDS-sliced image too large");
  X.task.assign((size_t)ntasks * 4, 0);
  std::vector<int32_t> nflag((size_t)ntasks, 0);
  std::vector<uint8_t> interior((size_t)ntasks, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < S; ++p) {
    for (int64_t t = tf[(size_t)p]; t < tf[(size_t)p + 1]; ++t) {
      const int64_t s0 = (t - tf[(size_t)p]) * TE;
      const int64_t eb = sb[(size_t)p] + s0, ec = std::min<int64_t>(TE, sne[(size_t)p] - s0);
      interior[(size_t)t] = (X.col[(size_t)eb] & 0x8000) ? 0 : 1;
      X.col[(size_t)eb] |= 0x8000;
      int32_t c = 0;
      for (int64_t e = eb; e < eb + ec; ++e) c += (X.col[(size_t)e] >> 15) & 1;
      nflag[(size_t)t] = c;
      X.task[(size_t)(4 * t)] = (int32_t)eb;
      X.task[(size_t)(4 * t + 1)] = (int32_t)ec;
      X.task[(size_t)(4 * t + 3)] = (int32_t)p;
    }
  }
  std::vector<int64_t> xint((size_t)ntasks + 1, 0);  // interior task starts before task t
  int64_t npieces = 0;
  for (int64_t t = 0; t < ntasks; ++t) {
    X.task[(size_t)(4 * t + 2)] = (int32_t)npieces;
    npieces += nflag[(size_t)t];
    xint[(size_t)t + 1] = xint[(size_t)t] + interior[(size_t)t];
  }
  if (npieces > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "CSR: LDS-sliced image too large");
  X.npieces = npieces;
  // ---- pieces of each row: first piece of each pair and the piece count
  std::vector<int32_t> rpieces((size_t)m, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; ++r) {
    int32_t c = 0;
    for (int64_t k = rpp[(size_t)r]; k < rpp[(size_t)r + 1]; ++k) {
      const int64_t p = pslc[(size_t)k], pos = ppos[(size_t)k], len = plen[(size_t)k];
      const int64_t rel = pos - sb[(size_t)p];
      const int64_t t = tf[(size_t)p] + rel / TE;
      const bool inside = interior[(size_t)t] && rel % TE != 0;
      // flags before pos: the pairs before it plus the interior task starts before it
      pidx[(size_t)k] = (int32_t)(pidx[(size_t)k] + xint[(size_t)t] + (inside ? 1 : 0));
      plen[(size_t)k] = (int32_t)(1 + (rel + len - 1) / TE - rel / TE);  // now: pieces of the pair
      c += plen[(size_t)k];
    }
    rpieces[(size_t)r] = c;
  }
  X.rowptr.assign((size_t)m + 1, 0);
  for (int64_t r = 0; r < m; ++r) X.rowptr[(size_t)r + 1] = X.rowptr[(size_t)r] + rpieces[(size_t)r];
  // ---- combine blocks
  X.cblk.clear();
  for (int64_t r = 0; r < m;) {
    const int64_t r0 = r;
    ++r;
    if (rpieces[(size_t)r0] <= comb_pieces)
      while (r < m && r - r0 < comb_rows && X.rowptr[(size_t)r + 1] - X.rowptr[(size_t)r0] <= comb_pieces)
        ++r;
    X.cblk.push_back((int32_t)r0);
    X.cblk.push_back((int32_t)r);
  }
  const int64_t nb = (int64_t)X.cblk.size() / 2;
  // ---- slot runs of each (block, slice) and the row-major places
  X.off.assign((size_t)(nb + 1) * S, -1);
  X.lpos.assign((size_t)std::max<int64_t>(npieces, 1), 0);  // never empty: the combine reads slot 0 when idle
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t r0 = X.cblk[(size_t)(2 * b)], r1 = X.cblk[(size_t)(2 * b + 1)];
    const int64_t base = X.rowptr[(size_t)r0];
    const bool big = X.rowptr[(size_t)r1] - base > comb_pieces;
    int32_t* o = X.off.data() + (size_t)b * S;
    for (int64_t r = r0; r < r1; ++r) {
      int64_t place = X.rowptr[(size_t)r] - base;
      for (int64_t k = rpp[(size_t)r]; k < rpp[(size_t)r + 1]; ++k) {
        const int64_t p = pslc[(size_t)k];
        if (o[p] < 0) o[p] = pidx[(size_t)k];
        for (int32_t i = 0; i < plen[(size_t)k]; ++i, ++place)
          if (!big) X.lpos[(size_t)pidx[(size_t)k] + i] = (uint16_t)place;
      }
    }
  }
  if (S > 0) {
    int32_t* last = X.off.data() + (size_t)nb * S;
    for (int64_t p = 0; p < S; ++p) {
      const int64_t t1 = tf[(size_t)p + 1];
      last[p] = t1 > tf[(size_t)p] ? X.task[(size_t)(4 * (t1 - 1) + 2)] + nflag[(size_t)(t1 - 1)]
                                   : (p > 0 ? last[p - 1] : 0);
    }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < S; ++p)
      for (int64_t b = nb - 1; b >= 0; --b)
        if (X.off[(size_t)(b * S + p)] < 0) X.off[(size_t)(b * S + p)] = X.off[(size_t)((b + 1) * S + p)];
  }
  // ---- workgroups: the tasks cut into nwg contiguous runs of about
  // nnz / nwg entries, each run cut again at slice boundaries
  X.units.clear();
  X.wg.assign((size_t)nwg + 1, 0);
  const int64_t target = std::max<int64_t>(1, (nnz + nwg - 1) / nwg);
  int64_t w = 0, acc = 0;
  for (int64_t p = 0; p < S; ++p) {
    int64_t t = tf[(size_t)p];
    const int64_t t1 = tf[(size_t)p + 1];
    while (t < t1) {
      if (acc >= target && w + 1 < nwg) {
        X.wg[(size_t)++w] = (int32_t)(X.units.size() / 3);
        acc = 0;
      }
      const int64_t start = t;
      while (t < t1 && (t == start || acc < target || w + 1 == nwg)) {
        acc += X.task[(size_t)(4 * t + 1)];
        ++t;
      }
      X.units.push_back((int32_t)p);
      X.units.push_back((int32_t)start);
      X.units.push_back((int32_t)t);
    }
  }
  for (int64_t k = w + 1; k <= nwg; ++k) X.wg[(size_t)k] = (int32_t)(X.units.size() / 3);
}

template void build_xlds<double>(XLdsHost<double>&, int64_t, int64_t, const int*, const int*,
                                 const double*, int64_t, int64_t, int64_t, int64_t, int);
template void build_xlds<float>(XLdsHost<float>&, int64_t, int64_t, const int*, const int*,
                                const float*, int64_t, int64_t, int64_t, int64_t, int);

}  // namespace rvvla
