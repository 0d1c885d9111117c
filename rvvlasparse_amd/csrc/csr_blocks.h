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

// Combine blocks of the LDS-sliced image: up to xl_comb_rows rows (one
// thread each) holding at most 8 partials a row on average (staged in LDS),
// or one row with more.  FP64 partials: 256 rows; 32-bit ones: 512 (longer
// runs of consecutive slots for the slice pass to write, measured faster).
constexpr int xl_comb_rows(int acc_bytes) { return acc_bytes == 8 ? 256 : 512; }
constexpr int xl_comb_pairs(int acc_bytes) { return 8 * xl_comb_rows(acc_bytes); }

}  // namespace rvvla
