/* VHCC.h -- vector-packed entry stream with segmented scans
 * (reference include/VHCC.h:15-115, kernel src/rvv/spmv_vhcc.cpp:73-250).
 * The CSR-ordered entry stream is cut into vec_size-wide vectors (64 on
 * gfx950).  _veceor_ptr lists the vectors holding at least one row end;
 * for each of them _scan_mask holds n_masks = log2(vec_size) scan masks
 * (lane j adds lane j - 2^s in step s) and one write mask (lanes that end a
 * row), vec_size/8 bytes per mask.  _row_arr lists the rows in the order
 * their ends appear.  _thr_info[t] gives the vector range [start_vec,
 * end_vec) of partition t with its first indices into _veceor_ptr
 * (vbase), _row_arr (rbase) and _scan_mask (sbase). */
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>
#include "CRS.h"
#include "device_image.h"

namespace SparseMatrixLib {

struct vhcc_params {
  size_t num_threads;
  size_t num_panels;
  size_t tile_row;
  size_t tile_col;
  size_t vec_size;
  size_t n_masks;
  vhcc_params(size_t nt = 1, size_t np = 1, size_t row_tile = 512, size_t col_tile = 512,
              size_t vec = 64)
      : num_threads(nt), num_panels(np), tile_row(row_tile), tile_col(col_tile), vec_size(vec),
        n_masks((size_t)std::log2((double)vec)) {}
};

struct thr_info_t {
  int vbase;
  int rbase;
  int sbase;
  int last_row;
  int overflow_row;
  int start_vec;
  int end_vec;
  int panel_id;
  int merge_start;
  int merge_end;
};

typedef uint32_t index_t;

template <typename ValT>
class spMtxVHCC {
  size_t _num_rows = 0;
  size_t _num_cols = 0;
  size_t _num_entries = 0;
  size_t _num_vectors = 0;
  size_t _nnz_per_panel = 0;
  size_t _pad_entries = 0;
  size_t _extended_rows = 0;
  size_t _pad_rows = 0;
  std::vector<thr_info_t> _thr_info;
  std::vector<uint32_t> _veceor_ptr;
  std::vector<uint8_t> _scan_mask;
  std::vector<uint32_t> _row_arr;
  std::vector<uint32_t> _col_arr;
  std::vector<ValT> _vals_arr;
  vhcc_params params;

 public:
  mutable DeviceImage device;

  size_t get_n_masks() const { return params.n_masks; }
  int get_pad_rows() const { return (int)_pad_rows; }
  size_t get_num_rows() const { return _num_rows; }
  size_t get_num_cols() const { return _num_cols; }
  size_t get_num_entries() const { return _num_entries; }
  size_t get_num_vectors() const { return _num_vectors; }
  size_t get_pad_entries() const { return _pad_entries; }
  const thr_info_t* get_thr_info() const { return _thr_info.data(); }
  const uint32_t* get_veceor_ptr() const { return _veceor_ptr.data(); }
  const uint8_t* get_scan_mask() const { return _scan_mask.data(); }
  const uint32_t* get_row_arr() const { return _row_arr.data(); }
  const uint32_t* get_col_arr() const { return _col_arr.data(); }
  const ValT* get_vals_arr() const { return _vals_arr.data(); }
  vhcc_params get_params() const { return params; }
  std::vector<thr_info_t>& get_thr_info_reference() { return _thr_info; }
  std::vector<uint32_t>& get_veceor_reference() { return _veceor_ptr; }
  std::vector<uint8_t>& get_scan_mask_reference() { return _scan_mask; }
  std::vector<uint32_t>& get_row_arr_reference() { return _row_arr; }
  std::vector<uint32_t>& get_col_arr_reference() { return _col_arr; }
  std::vector<ValT>& get_vals_arr_reference() { return _vals_arr; }
  const std::vector<thr_info_t>& thr_info() const { return _thr_info; }
  const std::vector<uint32_t>& veceor() const { return _veceor_ptr; }
  const std::vector<uint8_t>& scan_mask() const { return _scan_mask; }
  const std::vector<uint32_t>& row_arr() const { return _row_arr; }
  const std::vector<uint32_t>& col_arr() const { return _col_arr; }
  const std::vector<ValT>& vals_arr() const { return _vals_arr; }

  void freeMem() {
    _thr_info.clear(); _veceor_ptr.clear(); _scan_mask.clear();
    _row_arr.clear(); _col_arr.clear(); _vals_arr.clear();
  }
  void set_params(const vhcc_params& p) { params = p; }
  void set_properties(size_t rows, size_t cols, size_t entries, size_t vectors,
                      size_t pad_entries, size_t pad_rows) {
    _num_rows = rows; _num_cols = cols; _num_entries = entries;
    _num_vectors = vectors; _pad_entries = pad_entries; _pad_rows = pad_rows;
  }
};

}  // namespace SparseMatrixLib
