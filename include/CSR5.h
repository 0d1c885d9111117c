/* CSR5.h -- tile-transposed CSR (reference include/CSR5.h:29-112).
 * The nnz stream is cut into tiles of omega lanes x sigma entries; inside a
 * tile, lane l owns original entries tile_base + l*sigma + i, stored at
 * tile_base + i*omega + l.  Per lane, descriptor word 0 holds (MSB first)
 * y_offset (_bit_y_offset bits: rows started in lower lanes), seg_offset
 * (_bit_seg_offset bits: lanes to the next lane that starts a row) and the
 * sigma row-start flags, continued in words 1.._tile_desc_len-1; word w of
 * lane l is at tile*omega*_tile_desc_len + w*omega + l.  _csr5_tile_pointer
 * [t] is the row of the tile's first entry, bit 31 set when the tile spans
 * empty rows, whose segment -> row map is in _csr5_tile_empty_offset.
 * Rows >= _tail_tile_start are multiplied as CSR from entry (_p-1)*omega*sigma.
 * inputCSR/asCSR5 are declared by the reference and defined by this library. */
#pragma once
#include <cstdint>
#include <type_traits>
#include <vector>
#include "convert.h"
#include "device_image.h"

namespace SparseMatrixLib {

#define CSR5_SUCCESS 0
#define CSR5_UNKNOWN_FORMAT -1
#define CSR5_UNSUPPORTED_OMEGA -2
#define CSR5_CONVERTION_FAILED -3
#define CSR5_UNSUPPORTED_SPMV -4
#define CSR5_UNSUPPORTED_VALUE_TYPE -5

#define CSR5_FORMAT_CSR 0
#define CSR5_FORMAT_CSR5 1

/* One tile lane per wavefront lane on gfx950. */
#define CSR5_GFX950_OMEGA 64
#define CSR5_CACHELINE 64

template <typename T, typename U>
std::common_type_t<T, U> iceil(T num, U den) { return (num + den - 1) / den; }

inline uint32_t remove_sign(uint32_t v) { return v & 0x7FFFFFFFu; }

template <class ValT>
class spMtxCSR5 {
 public:
  /* Copy a CRS matrix in (keeps _format = CSR). */
  int inputCSR(int m, int n, int nnz, int* csr_row_pointer_ptr, int* csr_column_index_ptr,
               ValT* csr_value_ptr);
  /* Build the tiles; params keys "omega" (64 only) and "sigma" (0 = choose
   * from the mean row length). */
  int asCSR5(const convertParams& params);

  int _m = 0;
  int _n = 0;
  int _nnz = 0;
  int _bit_y_offset = 0;
  int _bit_seg_offset = 0;
  int _tile_desc_len = 0;
  int _tail_tile_start = 0;
  int _p = 0;
  int _omega = 0;
  int _sigma = 0;

  std::vector<int> _csr_row_pointer;
  std::vector<uint32_t> _csr_column_index;
  std::vector<ValT> _csr_value;
  std::vector<uint32_t> _csr5_tile_pointer;
  std::vector<uint32_t> _csr5_tile_descriptor;
  int _num_offsets = 0;
  std::vector<uint32_t> _csr5_tile_empty_offset_pointer;
  std::vector<uint32_t> _csr5_tile_empty_offset;
  std::vector<ValT> _temp_calibrator;
  mutable DeviceImage device;

  int format() const { return _format; }
  /* Marks the tile arrays as the CSR5 layout when they were filled from a
   * device-side build (convert with "device" = 1) instead of asCSR5. */
  void setCSR5Format() { _format = CSR5_FORMAT_CSR5; }
  void freeMem() {
    _csr_row_pointer.clear(); _csr_column_index.clear(); _csr_value.clear();
    _csr5_tile_pointer.clear(); _csr5_tile_descriptor.clear();
    _csr5_tile_empty_offset_pointer.clear(); _csr5_tile_empty_offset.clear();
    _temp_calibrator.clear();
  }

 private:
  int _format = CSR5_FORMAT_CSR;
};

}  // namespace SparseMatrixLib
