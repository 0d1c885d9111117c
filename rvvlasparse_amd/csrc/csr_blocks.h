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

// Combine blocks of the LDS-sliced image: up to kXlCombRows rows holding
// at most kXlCombPairs partials (staged in LDS), or one row with more.
constexpr int kXlCombRows = 256;
constexpr int kXlCombPairs = 2048;

}  // namespace rvvla
