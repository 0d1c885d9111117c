// hcsr_image.cpp -- the device image of an HCSR matrix.
//
// Input: the reference layout (reference include/HCSR.h fields: block_ptr,
// row_offset with the kind in bits 31:30, row_ptr, col_idx, values; the
// walk follows reference src/scalar/spmv_mtx_hcsr.cpp:52-76: CSR-kind
// blocks by their local row pointers, COO-kind blocks by one local row per
// entry), already validated by the caller.
//
// Output: work for one wavefront per slice of 64 consecutive rows of a
// block-row (a block-row of R rows has ceil(R / 64) slices; lane l owns
// the slice's row l).  For each block that holds entries of the slice:
//  * dense run: when the block's entries of the slice fill at least a
//    quarter of the rows-by-width rectangle (width = the most entries any of the
//    slice's rows has in the block), they are laid SELL-fashion: slot row j
//    of the run holds each lane's j-th entry at slot j * 64 + lane, with the
//    16-bit column inside the block and 0xFFFF marking an empty slot, and
//    each slot row carries the block's first column.  The runs of a slice
//    are consecutive slot rows.  (With C = 65536 a run holding column
//    0xFFFF goes to the tail instead.)
//  * tail: the other entries of the slice, as (32-bit column, 8-bit lane,
//    value), ordered by lane (stable: block order, then the block's own
//    order), padded to a whole number of 64-entry chunks.
// The slot rows of a slice's runs and its tail chunks are its work rows;
// a slice with more than kHcUnitRows of them is cut into several units
// (slot rows first, then tail chunks, in order), each of which
// leaves 64 partial sums in a buffer summed in unit order by a second pass.
// Rows whose slice has no entries still get a unit (it writes beta * y).
// The image ends with one empty slot row.
#include <omp.h>

#include <algorithm>

#include "internal.h"

namespace rvvla {

namespace {

template <class V>
struct BrowImage {
  std::vector<int32_t> unit, split, dxoff;
  std::vector<uint16_t> dcol;
  std::vector<V> dval;
  std::vector<int32_t> tcol;
  std::vector<uint8_t> trow;
  std::vector<V> tval;
  int64_t nparts = 0;
};

struct Ent {
  int32_t gcol;
  uint16_t c16;
  uint8_t lane;
  int32_t bc;
  int64_t src;  // entry index in values / col_idx
};

template <class V>
void build_brow(BrowImage<V>& B, int64_t br, int64_t m, int64_t R, int64_t C, int64_t bn,
                const int32_t* block_ptr, const uint32_t* row_offset, const uint32_t* row_ptr,
                const uint16_t* col_idx, const V* values) {
  const int64_t r0 = br * R, rows = std::min<int64_t>(R, m - r0);
  const int64_t ns = (rows + 63) / 64;
  // the block-row's entries grouped by slice (counting sort), block order
  // and each block's own order kept inside a slice
  std::vector<int64_t> cnt((size_t)ns + 1, 0);
  auto walk = [&](auto&& emit) {
    for (int64_t bc = 0; bc < bn; ++bc) {
      const int64_t b = br * bn + bc;
      const int64_t e0 = block_ptr[b], e1 = block_ptr[b + 1];
      if (e1 == e0) continue;
      const uint32_t kind = row_offset[b] >> 30, off = row_offset[b] & 0x3FFFFFFFu;
      if (kind == 1) {
        const uint32_t* rp = row_ptr + off;
        for (int64_t lr = 0; lr < rows; ++lr)
          for (int64_t e = e0 + rp[lr]; e < e0 + (int64_t)rp[lr + 1]; ++e) emit(bc, lr, e);
      } else {
        for (int64_t e = e0; e < e1; ++e) emit(bc, (int64_t)row_ptr[off + (e - e0)], e);
      }
    }
  };
  walk([&](int64_t, int64_t lr, int64_t) { ++cnt[(size_t)(lr / 64) + 1]; });
  for (int64_t s = 0; s < ns; ++s) cnt[(size_t)s + 1] += cnt[(size_t)s];
  std::vector<Ent> ent((size_t)cnt[(size_t)ns]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    walk([&](int64_t bc, int64_t lr, int64_t e) {
      const uint16_t c16 = col_idx[e];
      ent[(size_t)pos[(size_t)(lr / 64)]++] =
          Ent{(int32_t)(bc * C + c16), c16, (uint8_t)(lr % 64), (int32_t)bc, e};
    });
  }
  std::vector<int64_t> lane_cnt(64);
  std::vector<Ent> tail;
  for (int64_t s = 0; s < ns; ++s) {
    const int64_t srow0 = r0 + s * 64, nrows = std::min<int64_t>(64, rows - s * 64);
    tail.clear();
    const int64_t s0 = (int64_t)B.dxoff.size();  // the slice's first slot row
    int64_t work = 0;
    for (int64_t g0 = cnt[(size_t)s]; g0 < cnt[(size_t)s + 1];) {
      const int32_t bc = ent[(size_t)g0].bc;
      int64_t g1 = g0;
      std::fill(lane_cnt.begin(), lane_cnt.end(), 0);
      bool has_ffff = false;
      while (g1 < cnt[(size_t)s + 1] && ent[(size_t)g1].bc == bc) {
        ++lane_cnt[ent[(size_t)g1].lane];
        has_ffff |= ent[(size_t)g1].c16 == 0xFFFF;
        ++g1;
      }
      const int64_t w = *std::max_element(lane_cnt.begin(), lane_cnt.end());
      if (w * nrows <= 4 * (g1 - g0) && !has_ffff) {
        const int64_t rb = (int64_t)B.dcol.size() / 64;
        B.dcol.resize(B.dcol.size() + (size_t)(w * 64), 0xFFFF);
        B.dval.resize(B.dval.size() + (size_t)(w * 64), V(0));
        std::fill(lane_cnt.begin(), lane_cnt.end(), 0);
        for (int64_t g = g0; g < g1; ++g) {
          const Ent& q = ent[(size_t)g];
          const size_t slot = (size_t)((rb + lane_cnt[q.lane]++) * 64 + q.lane);
          B.dcol[slot] = q.c16;
          B.dval[slot] = values[q.src];
        }
        B.dxoff.resize(B.dxoff.size() + (size_t)w, (int32_t)((int64_t)bc * C));
        work += w;
      } else {
        tail.insert(tail.end(), ent.begin() + g0, ent.begin() + g1);
      }
      g0 = g1;
    }
    std::stable_sort(tail.begin(), tail.end(), [](const Ent& a, const Ent& b) { return a.lane < b.lane; });
    const int64_t te = (int64_t)B.tcol.size(), nt = (int64_t)tail.size();
    const int64_t chunks = (nt + 63) / 64;
    for (const Ent& q : tail) {
      B.tcol.push_back(q.gcol);
      B.trow.push_back(q.lane);
      B.tval.push_back(values[q.src]);
    }
    B.tcol.resize((size_t)(te + chunks * 64), 0);
    B.trow.resize((size_t)(te + chunks * 64), 0);
    B.tval.resize((size_t)(te + chunks * 64), V(0));
    const int64_t s1 = s0 + work;  // its slot rows are s0 .. s1
    work += chunks;
    // units: the work rows in order (runs, then tail chunks), at most
    // kHcUnitRows each; out = -1 for a slice done in one unit
    const int64_t nu = std::max<int64_t>(1, (work + kHcUnitRows - 1) / kHcUnitRows);
    const int64_t p0 = B.nparts;
    int64_t sr = s0;  // slot rows given out
    int64_t tc = 0;   // tail chunks given out
    for (int64_t u = 0; u < nu; ++u) {
      const int64_t take = std::min<int64_t>(kHcUnitRows, s1 - sr);
      const int64_t tk = std::min<int64_t>(kHcUnitRows - take, chunks - tc);
      const int64_t te0 = te + tc * 64, te1 = std::min(te + nt, te0 + tk * 64);
      B.unit.insert(B.unit.end(), {(int32_t)srow0, (int32_t)nrows, (int32_t)sr, (int32_t)(sr + take),
                                   (int32_t)te0, (int32_t)std::max(te0, te1),
                                   nu == 1 ? -1 : (int32_t)(p0 + u), 0});
      sr += take;
      tc += tk;
    }
    if (nu > 1) {
      B.split.insert(B.split.end(), {(int32_t)srow0, (int32_t)nrows, (int32_t)p0, (int32_t)(p0 + nu)});
      B.nparts += nu;
    }
  }
}

}  // namespace

template <class V>
void build_hcsr_image(HcsrImage<V>& im, int64_t m, int64_t n, int64_t R, int64_t C,
                      const int32_t* block_ptr, const uint32_t* row_offset,
                      const uint32_t* row_ptr, const uint16_t* col_idx, const V* values) {
  const int64_t bm = (m + R - 1) / R, bn = (n + C - 1) / C;
  std::vector<BrowImage<V>> B((size_t)bm);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t br = 0; br < bm; ++br)
    build_brow(B[(size_t)br], br, m, R, C, bn, block_ptr, row_offset, row_ptr, col_idx, values);
  // offsets of each block-row's pieces in the whole image
  struct Off { int64_t unit, split, drow, tent, part; };
  std::vector<Off> o((size_t)bm + 1, Off{0, 0, 0, 0, 0});
  for (int64_t br = 0; br < bm; ++br) {
    const BrowImage<V>& b = B[(size_t)br];
    const Off& p = o[(size_t)br];
    o[(size_t)br + 1] = Off{p.unit + (int64_t)b.unit.size(),
                            p.split + (int64_t)b.split.size(), p.drow + (int64_t)b.dcol.size() / 64,
                            p.tent + (int64_t)b.tcol.size(), p.part + b.nparts};
  }
  const Off& T = o[(size_t)bm];
  if (T.drow >= 2147483647LL || T.tent > 2147483647LL ||
      T.part > 2147483647LL || T.unit / kHcUnitInts > 2147483647LL)
    throw Error(RVVLA_ERR_LENGTH, "HCSR: device image exceeds the 32-bit offsets");
  // one empty slot row after the others: read by units without slot rows
  // in the kernel's grouped path
  im.dcol.assign((size_t)(T.drow + 1) * 64, 0xFFFF);
  im.dval.assign((size_t)(T.drow + 1) * 64, V(0));
  im.dxoff.assign((size_t)(T.drow + 1), 0);
  im.unit.resize((size_t)T.unit);
  im.split.resize((size_t)T.split);
  im.tcol.resize((size_t)T.tent);
  im.trow.resize((size_t)T.tent);
  im.tval.resize((size_t)T.tent);
  im.nunits = T.unit / kHcUnitInts;
  im.nsplit = T.split / 4;
  im.nparts = T.part;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t br = 0; br < bm; ++br) {
    const BrowImage<V>& b = B[(size_t)br];
    const Off& p = o[(size_t)br];
    for (size_t i = 0; i < b.unit.size(); i += kHcUnitInts) {
      int32_t* u = im.unit.data() + p.unit + i;
      std::copy(b.unit.begin() + (int64_t)i, b.unit.begin() + (int64_t)i + kHcUnitInts, u);
      u[2] += (int32_t)p.drow;
      u[3] += (int32_t)p.drow;
      u[4] += (int32_t)p.tent;
      u[5] += (int32_t)p.tent;
      if (u[6] >= 0) u[6] += (int32_t)p.part;
    }
    std::copy(b.dxoff.begin(), b.dxoff.end(), im.dxoff.begin() + p.drow);
    for (size_t i = 0; i < b.split.size(); i += 4) {
      int32_t* s = im.split.data() + p.split + i;
      std::copy(b.split.begin() + (int64_t)i, b.split.begin() + (int64_t)i + 4, s);
      s[2] += (int32_t)p.part;
      s[3] += (int32_t)p.part;
    }
    std::copy(b.dcol.begin(), b.dcol.end(), im.dcol.begin() + p.drow * 64);
    std::copy(b.dval.begin(), b.dval.end(), im.dval.begin() + p.drow * 64);
    std::copy(b.tcol.begin(), b.tcol.end(), im.tcol.begin() + p.tent);
    std::copy(b.trow.begin(), b.trow.end(), im.trow.begin() + p.tent);
    std::copy(b.tval.begin(), b.tval.end(), im.tval.begin() + p.tent);
  }
}

template void build_hcsr_image<double>(HcsrImage<double>&, int64_t, int64_t, int64_t, int64_t,
                                       const int32_t*, const uint32_t*, const uint32_t*,
                                       const uint16_t*, const double*);
template void build_hcsr_image<float>(HcsrImage<float>&, int64_t, int64_t, int64_t, int64_t,
                                      const int32_t*, const uint32_t*, const uint32_t*,
                                      const uint16_t*, const float*);

}  // namespace rvvla
