// csr_blocks.h -- row-block limits of the adaptive CSR kernel (shared by the
// host builders and the kernels).
#pragma once

namespace rvvla {

// One workgroup (kWG threads) per row block: either one row longer than
// kLongRow entries, or up to kStreamRows rows holding at most kStreamCap
// entries in total.
constexpr int kWG = 256;
constexpr int kStreamCap = 2048;
constexpr int kStreamRows = 256;
constexpr int kLongRow = 1024;

// LDS-sliced image (csr_xlds.cpp): x slices of kXlLdsBytes in LDS (16384
// doubles or 32768 floats), tasks of kXlTask slice-major entries (one
// wavefront each, 8 entries a lane).  The combine runs workgroups of cwg
// threads (256, 512 or 1024; kXlCombWG by default) over blocks of at most
// xl_comb_rows(cwg) rows holding at most xl_comb_pieces(acc, cwg) partials
// (64 bytes of LDS a thread), or one row with more.
constexpr int kXlLdsBytes = 131072;
constexpr int kXlTask = 512;
constexpr int xl_slice_cols(int io_bytes) { return kXlLdsBytes / io_bytes; }
constexpr int kXlCombWG = 1024;
constexpr int xl_comb_rows(int cwg) { return 2 * cwg; }
constexpr int xl_comb_pieces(int acc_bytes, int cwg) { return 64 * cwg / acc_bytes; }

// HCSR image (hcsr_image.cpp): ints per unit record (row0, nrows, first
// and end slot row, first and end tail entry, partial slot or -1, 0) and
// the most work rows (64 slots each) one unit takes.
constexpr int kHcUnitInts = 8;
constexpr int kHcUnitRows = 64;

}  // namespace rvvla
