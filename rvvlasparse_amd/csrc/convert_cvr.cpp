// convert_cvr.cpp -- CRS -> CVR host builder.
//
// The reference ships no CVR builder; the layout is the one its scalar
// kernel walks (reference src/scalar/spmv_mtx_cvr.cpp:32-66), built by the
// method of Xie et al. (CGO 2018): rows are split into Nthreads streams of
// contiguous rows with balanced entry counts; inside a stream SIMD_LEN lanes
// each take the next non-empty row whenever their current row is used up,
// so pack k holds one entry per lane.  Rows longer than cvr_segment_cap
// entries are taken as several segments of at most that many entries, each
// recorded under the row (the records add into y).  A lane that switches rows before
// pack k produces the record (k*SIMD_LEN + lane, finished row); a lane with
// nothing left holds a zero-valued pad entry (column of the stream's last
// entry).  Stream t's packs start at writeback_pos_record[t*(SIMD_LEN+2)],
// its records at writeback_pos_record[t*(SIMD_LEN+2)+1]; the rows still
// open when the stream ends are at +2 .. +SIMD_LEN+1 (-1: the lane holds no
// row; a single-stream layout puts the stream's last row there instead so
// the reference kernel, which adds every lane, stays in bounds).  Each
// stream's records end with a sentinel pair (end position, -1).
#include <omp.h>

#include <algorithm>

#include "internal.h"

namespace rvvla {

using namespace SparseMatrixLib;

int default_parts(int64_t m, int64_t nnz, int64_t items_per_part) {
  const int64_t p = (nnz + items_per_part - 1) / items_per_part;
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(m, 1), p));
}

std::vector<int64_t> balanced_row_parts(int64_t m, const int* rst, int64_t parts) {
  std::vector<int64_t> b((size_t)parts + 1, m);
  b[0] = 0;
  const int64_t nnz = m > 0 ? rst[m] : 0;
  for (int64_t t = 1; t < parts; ++t) {
    const int64_t target = (nnz * t + parts - 1) / parts;
    b[(size_t)t] = std::lower_bound(rst, rst + m + 1, (int)std::min<int64_t>(target, nnz)) - rst;
    b[(size_t)t] = std::max(b[(size_t)t], b[(size_t)t - 1]);
  }
  return b;
}

int64_t cvr_segment_cap(int64_t nnz, int64_t nstreams) {
  const int64_t per_lane = (nnz + nstreams * 64 - 1) / std::max<int64_t>(1, nstreams * 64);
  return std::max<int64_t>(64, per_lane);
}

template <class V>
void build_cvr(spMtxCVR<V>& d, const spMtxCRS<V>& s, int simd_len, int nstreams) {
  if (simd_len < 1 || simd_len > 64) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: SIMD_LEN in [1, 64]");
  if (nstreams < 1) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: at least one stream");
  const int64_t m = (int64_t)s.m, L = simd_len;
  // Rows longer than the cap (about one stream's share of entries per
  // lane) are cut into segments of cap entries, which the lanes take like
  // rows and whose records all name the row: the reference kernel adds
  // every record into y, so a row may be recorded once per segment.
  // Without the cut one lane would run a whole long row while the others
  // idle on padding.
  const int64_t cap = cvr_segment_cap((int64_t)s.nz, nstreams);
  std::vector<int> vrst{0};
  std::vector<int32_t> vrow;
  for (int64_t r = 0; r < m; ++r)
    for (int64_t e = s.Rst[r]; e < s.Rst[r + 1]; e += cap) {
      vrst.push_back((int)std::min<int64_t>(e + cap, s.Rst[r + 1]));
      vrow.push_back((int32_t)r);
    }
  const int64_t nv = (int64_t)vrow.size();
  const std::vector<int64_t> rb = balanced_row_parts(nv, vrst.data(), nstreams);
  // A stream lasts at least as many packs as its longest segment: refuse
  // layouts whose lower bound on packed slots exceeds 16x the entries (plus
  // 2^20) -- many more streams than entries
  {
    int64_t bound = 0;
    for (int t = 0; t < nstreams; ++t) {
      int64_t longest = 0;
      for (int64_t v = rb[(size_t)t]; v < rb[(size_t)t + 1]; ++v)
        longest = std::max<int64_t>(longest, vrst[(size_t)v + 1] - vrst[(size_t)v]);
      const int64_t nz_t = vrst[(size_t)rb[(size_t)t + 1]] - vrst[(size_t)rb[(size_t)t]];
      bound += std::max<int64_t>((nz_t + L - 1) / L, longest) * L;
    }
    if (bound > 16 * (int64_t)s.nz + (int64_t(1) << 20) || bound > 2147483647LL)
      throw Error(RVVLA_ERR_UNSUPPORTED_SPMV,
                  "CVR: too many streams for the entries (at least " + std::to_string(bound) +
                      " packed slots for " + std::to_string(s.nz) + " entries)");
  }
  // pass 1 (parallel): each stream packed on its own
  struct Stream {
    std::vector<V> vals;
    std::vector<uint32_t> cols;
    std::vector<int> rec;  // (pos relative to the stream, row) pairs
    std::vector<int> tail;
  };
  std::vector<Stream> st((size_t)nstreams);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < nstreams; ++t) {
    Stream& S = st[(size_t)t];
    int64_t next = rb[(size_t)t];
    const int64_t end = rb[(size_t)t + 1];
    std::vector<int64_t> seg((size_t)L, -1), pos((size_t)L, 0), stop((size_t)L, 0);
    uint32_t pad_col = 0;
    if (end > rb[(size_t)t]) pad_col = (uint32_t)s.Col[vrst[(size_t)end] - 1];
    auto take = [&](int64_t lane) {
      if (next >= end) {
        seg[(size_t)lane] = -1;
        return;
      }
      seg[(size_t)lane] = next;
      pos[(size_t)lane] = vrst[(size_t)next];
      stop[(size_t)lane] = vrst[(size_t)next + 1];
      ++next;
    };
    for (int64_t l = 0; l < L; ++l) take(l);
    for (int64_t k = 0;; ++k) {
      bool any = false;
      for (int64_t l = 0; l < L; ++l) {
        if (seg[(size_t)l] >= 0 && pos[(size_t)l] == stop[(size_t)l]) {
          S.rec.push_back((int)(k * L + l));
          S.rec.push_back(vrow[(size_t)seg[(size_t)l]]);
          take(l);
        }
        any |= seg[(size_t)l] >= 0;
      }
      if (!any) break;
      for (int64_t l = 0; l < L; ++l) {
        if (seg[(size_t)l] >= 0) {
          S.vals.push_back(s.Val[pos[(size_t)l]]);
          S.cols.push_back((uint32_t)s.Col[pos[(size_t)l]]);
          ++pos[(size_t)l];
        } else {
          S.vals.push_back(V(0));
          S.cols.push_back(pad_col);
        }
      }
    }
    // every lane ended idle: the last records flushed all rows
    S.tail.assign((size_t)L, -1);
  }
  // pass 2: concatenate; the last record of each stream may point at the pack
  // just past its end (the flush after the final pack)
  d = spMtxCVR<V>();
  d.SIMD_LEN = (size_t)L;
  d.numRows = (size_t)m;
  d.numCols = s.n;
  d.Nthreads = nstreams;
  d.Nthreads_avail = nstreams;
  d.writeback_pos_record.assign((size_t)nstreams * (L + 2), -1);
  int64_t items = 0, recs = 0;
  for (int t = 0; t < nstreams; ++t) {
    const Stream& S = st[(size_t)t];
    int* wb = d.writeback_pos_record.data() + (size_t)t * (L + 2);
    wb[0] = (int)(items / L);
    wb[1] = (int)(recs / 2);
    for (int64_t l = 0; l < L; ++l) wb[2 + l] = S.tail[(size_t)l];
    for (size_t r = 0; r < S.rec.size(); r += 2) {
      d.vPack_vec_record.push_back((int)(items + S.rec[r]));
      d.vPack_vec_record.push_back(S.rec[r + 1]);
    }
    recs += (int64_t)S.rec.size();
    items += (int64_t)S.vals.size();
    if (items > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "CVR: more than 2^31 packed entries");
    d.vPack_vec_vals.insert(d.vPack_vec_vals.end(), S.vals.begin(), S.vals.end());
    d.vPack_vec_cols.insert(d.vPack_vec_cols.end(), S.cols.begin(), S.cols.end());
  }
  // the reference's kernel reads a final sentinel pair
  d.vPack_vec_record.push_back((int)items);
  d.vPack_vec_record.push_back(-1);
  d.nItems = (size_t)items;
  if (nstreams == 1) {
    // single stream: give idle lanes a valid row (they add alpha * 0 there)
    int last = -1;
    for (int64_t r = m - 1; r >= 0; --r)
      if (s.Rst[r + 1] > s.Rst[r]) { last = (int)r; break; }
    for (int64_t l = 0; l < L; ++l)
      if (d.writeback_pos_record[2 + l] < 0) d.writeback_pos_record[2 + l] = last < 0 ? 0 : last;
  }
}

template void build_cvr<double>(spMtxCVR<double>&, const spMtxCRS<double>&, int, int);
template void build_cvr<float>(spMtxCVR<float>&, const spMtxCRS<float>&, int, int);

}  // namespace rvvla
