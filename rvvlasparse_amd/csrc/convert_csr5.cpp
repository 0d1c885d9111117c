// convert_csr5.cpp -- CRS -> CSR5 host builder.
//
// The reference declares spMtxCSR5::inputCSR / asCSR5 (reference
// include/CSR5.h:61-65) without defining them; they are defined here.  The
// layout is the one the reference's vector kernel reads
// (reference src/rvv/spmv_mtx_csr5.cpp:85-296, tail :337-365), built by the
// method of Liu & Vinter (ICS 2015):
//  * p = ceil(nnz / (omega*sigma)) tiles; tiles 0..p-2 are full and stored
//    transposed (entry i of lane l at tile_base + i*omega + l, original
//    position tile_base + l*sigma + i); entries from (p-1)*omega*sigma on
//    stay in CSR order and rows from _tail_tile_start are multiplied as CSR.
//  * tile_ptr[t] = row holding entry t*omega*sigma (the last row starting at
//    or before it, never an empty one); bit 31 marks a tile whose row range
//    has empty rows strictly inside it; tile_ptr[p] = m.
//  * row-start flag (lane l, entry i) = the entry begins a row.  The flag of
//    a tile's very first entry only says whether tile_ptr[t] starts there and
//    is not counted: y_offset of lane l = flags in lanes < l (others), so the
//    k-th counted flag of a tile (k = 0, 1, ...) starts row tile_ptr[t]+1+k,
//    or tile_ptr[t]+1+_csr5_tile_empty_offset[empty_offset_pointer[t]+k] in
//    a tile with empty rows.
//  * seg_offset of lane l = number of lanes after l before the next lane
//    holding a flag (or before the tile end).
//  * descriptor word 0 = y_offset | seg_offset | flags 0.. (MSB first),
//    words 1.. = 32 further flags each; word w of lane l at
//    tile*omega*_tile_desc_len + w*omega + l.
#include <omp.h>

#include <algorithm>

#include "internal.h"

namespace SparseMatrixLib {

template <class ValT>
int spMtxCSR5<ValT>::inputCSR(int m, int n, int nnz, int* rp, int* ci, ValT* v) {
  if (m < 0 || n < 0 || nnz < 0 || (m > 0 && !rp) || (nnz > 0 && (!ci || !v)))
    return CSR5_CONVERTION_FAILED;
  _m = m;
  _n = n;
  _nnz = nnz;
  _csr_row_pointer.assign(rp, rp + m + 1);
  _csr_column_index.assign(ci, ci + nnz);
  _csr_value.assign(v, v + nnz);
  _format = CSR5_FORMAT_CSR;
  return CSR5_SUCCESS;
}

static int ceil_log2(int64_t v) {
  int b = 0;
  while ((int64_t(1) << b) < v) ++b;
  return b;
}

template <class ValT>
int spMtxCSR5<ValT>::asCSR5(const convertParams& params) {
  if (_format == CSR5_FORMAT_CSR5) return CSR5_SUCCESS;
  const int omega = rvvla::param_int(params, "omega", CSR5_GFX950_OMEGA);
  if (omega != CSR5_GFX950_OMEGA) return CSR5_UNSUPPORTED_OMEGA;
  int sigma = rvvla::param_int(params, "sigma", 0);
  if (sigma <= 0) sigma = rvvla::default_csr5_sigma(_m, _nnz);
  if (sigma > 64) return CSR5_CONVERTION_FAILED;
  const int64_t m = _m, nnz = _nnz, W = (int64_t)omega * sigma;
  const int* rst = _csr_row_pointer.data();
  _omega = omega;
  _sigma = sigma;
  _p = nnz > 0 ? (int)((nnz + W - 1) / W) : 1;
  _bit_y_offset = ceil_log2(W);
  _bit_seg_offset = ceil_log2(omega);
  _tile_desc_len = (_bit_y_offset + _bit_seg_offset + sigma + 31) / 32;
  const int64_t ntiles = _p - 1;  // full tiles
  const int dl = _tile_desc_len, f0 = 32 - _bit_y_offset - _bit_seg_offset;

  // tile_ptr: row of each tile's first entry (m past the end)
  _csr5_tile_pointer.assign((size_t)_p + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t <= _p; ++t) {
    const int64_t e = t * W;
    _csr5_tile_pointer[(size_t)t] =
        e < nnz ? (uint32_t)(std::upper_bound(rst, rst + m + 1, (int)e) - rst - 1) : (uint32_t)m;
  }
  // empty rows: prefix count, tile flags
  std::vector<int64_t> nempty((size_t)m + 1, 0);
  for (int64_t r = 0; r < m; ++r) nempty[(size_t)r + 1] = nempty[(size_t)r] + (rst[r] == rst[r + 1]);
  std::vector<int64_t> cnt((size_t)std::max<int64_t>(ntiles, 0) + 1, 0);
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t r0 = _csr5_tile_pointer[(size_t)t], r1 = _csr5_tile_pointer[(size_t)t + 1];
    if (r1 > r0 + 1 && nempty[(size_t)r1] - nempty[(size_t)r0 + 1] > 0) {
      _csr5_tile_pointer[(size_t)t] |= 0x80000000u;
      // counted flags of the tile = rows starting strictly inside it
      const int64_t b = t * W;
      const int64_t lo = std::upper_bound(rst, rst + m + 1, (int)b) - rst;       // first row start > b
      const int64_t hi = std::lower_bound(rst, rst + m + 1, (int)(b + W)) - rst; // first row start >= b+W
      int64_t c = 0;
      for (int64_t r = lo; r < hi && r < m; ++r) c += rst[r + 1] > rst[r];
      cnt[(size_t)t + 1] = c;
    }
  }
  _csr5_tile_empty_offset_pointer.assign((size_t)_p + 1, 0);
  for (int64_t t = 0; t < ntiles; ++t) cnt[(size_t)t + 1] += cnt[(size_t)t];
  for (int64_t t = 0; t < ntiles; ++t) _csr5_tile_empty_offset_pointer[(size_t)t] = (uint32_t)cnt[(size_t)t];
  if (ntiles > 0) _csr5_tile_empty_offset_pointer[(size_t)ntiles] = (uint32_t)cnt[(size_t)ntiles];
  _num_offsets = ntiles > 0 ? (int)cnt[(size_t)ntiles] : 0;
  _csr5_tile_empty_offset.assign((size_t)_num_offsets, 0);
  _csr5_tile_descriptor.assign((size_t)std::max<int64_t>(ntiles, 0) * omega * dl, 0u);

#pragma omp parallel
  {
    std::vector<uint64_t> flags((size_t)omega * 2);
#pragma omp for schedule(dynamic, 16)
    for (int64_t t = 0; t < ntiles; ++t) {
      const int64_t base = t * W;
      const bool has_empty = (_csr5_tile_pointer[(size_t)t] >> 31) != 0;
      const int64_t row0 = _csr5_tile_pointer[(size_t)t] & 0x7FFFFFFFu;
      int64_t r = row0;
      uint32_t ord = 0;
      size_t eo = has_empty ? _csr5_tile_empty_offset_pointer[(size_t)t] : 0;
      std::vector<uint32_t> yoff((size_t)omega);
      std::vector<uint8_t> any((size_t)omega, 0);
      std::fill(flags.begin(), flags.end(), 0);
      for (int l = 0; l < omega; ++l) {
        yoff[(size_t)l] = ord;
        for (int i = 0; i < sigma; ++i) {
          const int64_t e = base + (int64_t)l * sigma + i;
          while (rst[r + 1] <= e) ++r;
          if (rst[r] != e) continue;
          flags[(size_t)l * 2 + (i >> 6)] |= uint64_t(1) << (i & 63);
          if (l == 0 && i == 0) continue;  // tile_ptr[t] itself: not counted
          any[(size_t)l] = 1;
          if (has_empty) _csr5_tile_empty_offset[eo++] = (uint32_t)(r - row0 - 1);
          ++ord;
        }
      }
      uint32_t* d = _csr5_tile_descriptor.data() + (size_t)t * omega * dl;
      int next = omega;
      for (int l = omega - 1; l >= 0; --l) {
        const uint32_t seg = (uint32_t)(next - l - 1);
        if (any[(size_t)l]) next = l;
        uint32_t w0 = (yoff[(size_t)l] << (32 - _bit_y_offset)) |
                      (seg << (32 - _bit_y_offset - _bit_seg_offset));
        for (int i = 0; i < sigma; ++i) {
          if (!((flags[(size_t)l * 2 + (i >> 6)] >> (i & 63)) & 1)) continue;
          if (i < f0) w0 |= 1u << (31 - (32 - f0) - i);
          else d[(size_t)(1 + (i - f0) / 32) * omega + l] |= 1u << (31 - (i - f0) % 32);
        }
        d[l] = w0;
      }
    }
  }
  // transpose the full tiles
  if (ntiles > 0) {
    std::vector<uint32_t> c2(_csr_column_index);
    std::vector<ValT> v2(_csr_value);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < ntiles; ++t)
      for (int l = 0; l < omega; ++l)
        for (int i = 0; i < sigma; ++i) {
          const size_t src = (size_t)(t * W + (int64_t)l * sigma + i);
          const size_t dst = (size_t)(t * W + (int64_t)i * omega + l);
          c2[dst] = _csr_column_index[src];
          v2[dst] = _csr_value[src];
        }
    _csr_column_index.swap(c2);
    _csr_value.swap(v2);
  }
  _tail_tile_start = (int)(_csr5_tile_pointer[(size_t)_p - 1] & 0x7FFFFFFFu);
  _temp_calibrator.assign((size_t)omp_get_max_threads() * (CSR5_CACHELINE / sizeof(ValT)), ValT(0));
  _format = CSR5_FORMAT_CSR5;
  return CSR5_SUCCESS;
}

template class spMtxCSR5<double>;
template class spMtxCSR5<float>;

}  // namespace SparseMatrixLib

namespace rvvla {

using namespace SparseMatrixLib;

// sigma = 8: measured on MI355X against 4 and 16 it is the fastest or
// within 1 % on the 5-point Laplacian (73 us against 85 us at 16, FP64),
// the power-law matrix and 16 uniform entries per row
int default_csr5_sigma(int64_t m, int64_t nnz) {
  (void)m;
  (void)nnz;
  return 8;
}

template <class V>
void build_csr5(spMtxCSR5<V>& d, const spMtxCRS<V>& s, int omega, int sigma) {
  d.freeMem();
  int rc = d.inputCSR((int)s.m, (int)s.n, (int)s.nz, s.Rst, s.Col, s.Val);
  if (rc) throw Error(RVVLA_ERR_CONVERSION, "CSR5: inputCSR failed");
  convertParams P;
  P.param["omega"].i = omega;
  P.param["sigma"].i = sigma;
  rc = d.asCSR5(P);
  if (rc == CSR5_UNSUPPORTED_OMEGA) throw Error(RVVLA_ERR_UNSUPPORTED_OMEGA, "CSR5: omega must be 64");
  if (rc) throw Error(RVVLA_ERR_CONVERSION, "CSR5: sigma must be in [1, 64]");
}

template void build_csr5<double>(spMtxCSR5<double>&, const spMtxCRS<double>&, int, int);
template void build_csr5<float>(spMtxCSR5<float>&, const spMtxCRS<float>&, int, int);

}  // namespace rvvla
