// internal.h -- shared declarations of librvvla_hip.so (not installed).
#pragma once
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "csr_blocks.h"
#include "rvvla_hip.h"
#include "sparse_matrix.h"

namespace rvvla {

// Error carrying an RVVLA_* code; converted to the return value at the ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

// Last error text of the calling thread (rvvla_strerror appends it).
void set_last_error(const std::string& s);
const std::string& last_error();

// Validates CSR arrays (shape, monotone row pointers, columns in range);
// throws Error(RVVLA_ERR_INVALID_ARG).
void check_csr_public(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col);
// Key of a matrix's row-length histogram (autotune.cpp): rows in
// power-of-two length buckets rounded to 1/20, the max/mean bucket and the
// right-hand-side count rounded up to a power of two.
std::string row_length_key(int64_t m, const int* rst, int k);

// convertParams from "k=v,k=v" text.
SparseMatrixLib::convertParams parse_params(const char* text);
int param_int(const SparseMatrixLib::convertParams& p, const char* key, int dflt);
double param_double(const SparseMatrixLib::convertParams& p, const char* key, double dflt);

// ---- host converters (CRS -> reference container fields) -----------------
template <class V>
void build_hcsr(SparseMatrixLib::spMtxHCSR<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
                size_t R, size_t C, double p);
template <class V>
void build_sell(SparseMatrixLib::spMtxSELL_C_Sigma<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
                int C, int sigma);
template <class V>
void build_csr5(SparseMatrixLib::spMtxCSR5<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
                int omega, int sigma);
template <class V>
void build_lav(SparseMatrixLib::spMtxLAV<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
               int lanes, double T, int segment_len);
template <class V>
void build_cvr(SparseMatrixLib::spMtxCVR<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
               int simd_len, int nstreams);
template <class V>
void build_vhcc(SparseMatrixLib::spMtxVHCC<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
                int vec_size, int nparts);
template <class V>
void build_vnec(SparseMatrixLib::spMtxVNEC<V>& d, const SparseMatrixLib::spMtxCRS<V>& s,
                int nparts);
// The IRD variant selector of a VNEC whose partitions, coordinates and ranks
// are filled (IRD_mat, vnec_t).
template <class V>
void vnec_select_variant(SparseMatrixLib::spMtxVNEC<V>& d, const SparseMatrixLib::spMtxCRS<V>& s);

// Default partition counts and parameters chosen for gfx950 (256 CUs).
int default_sell_sigma(int64_t m, int64_t nnz, const int* rst);
// SELL-64 slots when all rows are sorted by length (sigma = m).
int64_t sorted_sell_slots(int64_t m, const int* rst);
int default_csr5_sigma(int64_t m, int64_t nnz);
int default_parts(int64_t m, int64_t nnz, int64_t items_per_part);
// Default work per wavefront (measured on the 4.2 M-row Laplacian): CVR
// streams of about 4096 entries, VHCC partitions of about 1024.
constexpr int64_t kCvrStreamEntries = 4096;
constexpr int64_t kVhccPartEntries = 1024;
// Rows [b[t], b[t+1]) of part t: contiguous, about nnz/parts entries each.
std::vector<int64_t> balanced_row_parts(int64_t m, const int* rst, int64_t parts);
// CVR: the most entries a lane takes as one row; longer rows are cut into
// segments of this many (at least 64, about one stream's share per lane).
int64_t cvr_segment_cap(int64_t nnz, int64_t nstreams);

// The device image of an HCSR matrix (hcsr_image.cpp) from the validated
// reference layout.
template <class V>
struct HcsrImage {
  std::vector<int32_t> unit, split, dxoff, tcol;
  std::vector<uint16_t> dcol;
  std::vector<uint8_t> trow;
  std::vector<V> dval, tval;
  int64_t nunits = 0, nsplit = 0, nparts = 0;
};
template <class V>
void build_hcsr_image(HcsrImage<V>& im, int64_t m, int64_t n, int64_t R, int64_t C,
                      const int32_t* block_ptr, const uint32_t* row_offset,
                      const uint32_t* row_ptr, const uint16_t* col_idx, const V* values);

// ---- device matrices ----------------------------------------------------
// Upload of each format's host container; value arrays are given as V
// (double for F64, float otherwise) and narrowed for BF16/F16.
// xslices: column slices of the sliced image (-1 = chosen from the matrix
// and the device, 0 or 1 = plain CSR only, 2..8 = that many).  xlds: the
// LDS-sliced image (-1 = chosen from the matrix, 0 = off, 1 = on; it takes
// the place of the sliced image).  sell: the natural-order SELL-64 image
// for matrices with even row lengths (-1 = chosen from the matrix, 0 = off,
// 1 = on; only where no sliced image is used).  xlcomb: threads of the
// LDS-sliced image's combine workgroups (256, 512 or 1024; 0 = default).
template <class V>
rvvla_mat_t upload_csr(int dtype, const SparseMatrixLib::spMtxCRS<V>& a, int xslices = -1,
                       int xlds = -1, int sell = -1, int xlcomb = 0);
template <class V>
rvvla_mat_t upload_hcsr(int dtype, const SparseMatrixLib::spMtxHCSR<V>& a);
template <class V>
rvvla_mat_t upload_sell(int dtype, const SparseMatrixLib::spMtxSELL_C_Sigma<V>& a);
template <class V>
rvvla_mat_t upload_csr5(int dtype, const SparseMatrixLib::spMtxCSR5<V>& a);
template <class V>
rvvla_mat_t upload_lav(int dtype, const SparseMatrixLib::spMtxLAV<V>& a);
template <class V>
rvvla_mat_t upload_cvr(int dtype, const SparseMatrixLib::spMtxCVR<V>& a, int64_t nnz);
template <class V>
rvvla_mat_t upload_vhcc(int dtype, const SparseMatrixLib::spMtxVHCC<V>& a);
template <class V>
rvvla_mat_t upload_vnec(int dtype, const SparseMatrixLib::spMtxVNEC<V>& a,
                        const SparseMatrixLib::spMtxCRS<V>& src);

// The transposed product of a matrix in another format runs on a CSR image
// of A^T built, on its first use, from the CSR the matrix was converted
// from: attach_transpose_source keeps `src` (shared) for that.
template <class V>
void attach_transpose_source(rvvla_mat_s& M, std::shared_ptr<const SparseMatrixLib::spMtxCRS<V>> src);

// Adaptive CSR row blocks over rows [r0, r1) as (first, end) pairs: one
// row of more than kLongRow entries, or up to kStreamRows rows holding at
// most kStreamCap entries; the long-row blocks first, longest first, then
// the others in row order.
std::vector<int32_t> csr_row_blocks(const int* rst, int64_t r0, int64_t r1);

// Column-sliced CSR image (csr_xslice.cpp).
template <class V>
struct XSliceHost {
  int P = 0;
  int64_t width = 0, nq = 0;
  std::vector<int32_t> rst, col, blk, base, pref;
  std::vector<V> val;
  std::vector<uint16_t> mask;
};
// P slices (2..16) on `xcd` XCDs (slices are grouped into phases of xcd).
template <class V>
void build_xslice(XSliceHost<V>& X, int64_t m, int64_t n, const int* rst, const int* col,
                  const V* val, int P, int xcd);
// LDS-sliced CSR image (csr_xlds.cpp): x cut into slices of W elements
// that fill one workgroup's LDS; entries slice-major with 16-bit columns,
// cut into tasks of kXlTask entries; one partial per piece (a (row, slice)
// pair's entries inside one task).
template <class V>
struct XLdsHost {
  int64_t W = 0, S = 0, npieces = 0, nent = 0;
  std::vector<uint16_t> col, lpos;
  std::vector<V> val;
  std::vector<int32_t> task, units, wg, rowptr, cblk, off;
};
// (row, slice) pairs of a CSR matrix for slices of W columns.
int64_t xlds_pairs(int64_t m, const int* rst, const int* col, int64_t W);
// nwg: the workgroups of the slice pass; combine blocks of at most
// comb_rows rows and comb_pieces partials (or one row with more); vpw:
// device values per 16 bytes (2 FP64, 4 FP32, 8 for 16-bit values)
template <class V>
void build_xlds(XLdsHost<V>& X, int64_t m, int64_t n, const int* rst, const int* col, const V* val,
                int64_t W, int64_t nwg, int64_t comb_rows, int64_t comb_pieces, int vpw);
// Slot of stream position `at` in the value array (values of a task
// permuted for contiguous 16-byte loads).
int64_t xl_value_slot(int64_t at, int vpw);
// Fraction of entries whose column is more than `window` away from the
// row's diagonal position r * n / m.
double far_fraction(int64_t m, int64_t n, const int* rst, const int* col, int64_t window);

// 16-bit value encodings (round to nearest even).
uint16_t f32_to_bf16(float f);
uint16_t f32_to_f16(float f);
float bf16_to_f32(uint16_t b);
float f16_to_f32(uint16_t b);

}  // namespace rvvla
