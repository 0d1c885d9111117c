// k_csr5.hip -- CSR5 SpMV for gfx950 (omega = 64: one tile lane per
// wavefront lane).
//
// Semantics follow the reference vector kernel (reference
// src/rvv/spmv_mtx_csr5.cpp:85-296 compute, :318-335 calibrate, :337-365
// tail): per tile, lane partial sums split at row-start flags, a segmented
// sum across lanes, and a CSR pass over the tail rows.  The reference fixes
// rows that cross tile boundaries with per-thread calibrators; here every
// row is finished by exactly one writer, so no atomics are needed and the
// result is deterministic:
//  * k_csr5_tiles (one wavefront per full tile): rows that start and end
//    inside the tile are written directly; the tile's leading segment
//    (before its first counted flag) goes to first_part[t], and the segment
//    opened by its last flag to last_part[t] / last_row[t].
//  * k_csr5_fixup (one thread per full tile, then one per tail row, then
//    one per empty row): the owner of each open segment adds the following
//    tiles' leading segments up to the tile where the row ends (that range
//    is fixed by the structure and found at upload), tail rows are summed
//    as CSR, empty rows get alpha*0 + beta*y.  Segments spanning more than
//    kCsr5Serial tiles or running into the tail, and tail rows longer than
//    kCsr5Serial entries, take a wavefront each (csr5_long_item) in extra
//    workgroups of the same launch.
//  * With 32-bit accumulation and sigma 8 or 16, each XCD walks a
//    contiguous run of tiles.
#include "kcommon.h"

namespace rvvla {

struct Csr5Args {
  const uint32_t* tile_ptr;
  const uint32_t* desc;
  const uint32_t* eptr;
  const uint32_t* eoff;
  const int32_t* rst;
  const uint32_t* col;
  const int32_t* empty_rows;
  const int32_t* chain;
  const int32_t* long_fix;
  int64_t ntiles, nempty, m, nlong;
  int sigma, dl, bit_y, bit_seg, tail_start, first_tail_row;
  int64_t tail_base;  // (p-1)*omega*sigma
};

// flag bit i (0-based) of one lane, from its descriptor words
__device__ __forceinline__ uint64_t csr5_flags(const uint32_t* w, int sigma, int f0) {
  uint64_t fl = 0;
  for (int i = 0; i < sigma; ++i) {
    uint32_t bit;
    if (i < f0) bit = (w[0] >> (31 - (32 - f0) - i)) & 1u;
    else bit = (w[1 + (i - f0) / 32] >> (31 - (i - f0) % 32)) & 1u;
    fl |= (uint64_t)bit << i;
  }
  return fl;
}

// SIG > 0: sigma known at compile time, one tile per 64-thread workgroup,
// and the x gathers are made in CSR order through LDS: the tile's columns
// are staged (transposed back to CSR order), each lane gathers x for 64
// consecutive entries at a time, and the values of x go back to the lane
// that owns the entry in the tile layout.  SIG = 0: any sigma, four tiles
// per workgroup, gathers in tile order.
template <class VT, int SIG, bool COMP>
__global__ __launch_bounds__(SIG > 0 ? 64 : kWG) void k_csr5_tiles(
    Csr5Args a, const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc* __restrict__ first_part,
    typename Val<VT>::acc* __restrict__ last_part, int32_t* __restrict__ last_row,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  using Ac = Acc<A, COMP>;
  constexpr int TPB = SIG > 0 ? 1 : kWG / 64;  // tiles per workgroup
  __shared__ A xb[SIG > 0 ? 64 * (SIG + 1) : 1];
  const int64_t t = SIG > 0 && sizeof(A) < 8 ? (int64_t)xcd_run_index((int)blockIdx.x, (int)gridDim.x) : (int64_t)blockIdx.x * TPB + (threadIdx.x >> 6);
  if (t >= a.ntiles) return;  // uniform per wavefront (per workgroup when SIG > 0)
  const int lane = threadIdx.x & 63;
  const int sigma = a.sigma, f0 = 32 - a.bit_y - a.bit_seg;
  const uint32_t tp = a.tile_ptr[t];
  const bool has_empty = (tp >> 31) != 0;
  const int64_t row0 = tp & 0x7FFFFFFFu;
  uint32_t w[3] = {0, 0, 0};
  const uint32_t* dp = a.desc + t * 64 * a.dl + lane;
  for (int k = 0; k < a.dl; ++k) w[k] = dp[k * 64];
  const uint32_t yoff = w[0] >> (32 - a.bit_y);
  uint64_t fl = csr5_flags(w, sigma, f0);
  const bool starts_here = lane == 0 && (fl & 1u);  // tile_ptr[t] begins at this tile
  if (lane == 0) fl &= ~uint64_t(1);
  const uint32_t ebase = has_empty ? a.eptr[t] : 0u;
  auto row_of = [&](uint32_t ord) -> int64_t {  // ord: 1-based counted-flag ordinal
    return has_empty ? row0 + 1 + (int64_t)a.eoff[ebase + ord - 1] : row0 + (int64_t)ord;
  };
  Ac head, cur;
  uint32_t nfl = 0;
  const int64_t base = t * 64 * (int64_t)sigma + lane;
  const uint32_t* __restrict__ colp = a.col;
  auto walk = [&](int i, A p) {  // the flag walk over lane step i
    if ((fl >> i) & 1u) {
      if (nfl == 0) {
        head = cur;
      } else {
        const int64_t r = row_of(yoff + nfl);
        y[r] = alpha * cur.value() + scaled_y(beta, y, r);
      }
      cur = Ac();
      ++nfl;
    }
    cur.add(p);
  };
  if constexpr (SIG > 0) {
    constexpr int W = sizeof(A) / 4;  // 32-bit words per LDS slot
    uint32_t* cbits = reinterpret_cast<uint32_t*>(xb);
    A v[SIG];
#pragma unroll
    for (int i = 0; i < SIG; ++i) {
      const int64_t e = base + (int64_t)i * 64;
      const uint32_t c = colp[e];
      v[i] = Val<VT>::get(val, e);
      cbits[(lane * (SIG + 1) + i) * W] = c;  // entry lane*SIG + i, padded row
    }
    __syncthreads();
    uint32_t cc[SIG];
#pragma unroll
    for (int q = 0; q < SIG; ++q) {
      const int j = q * 64 + lane;  // CSR order within the tile
      cc[q] = cbits[((j / SIG) * (SIG + 1) + (j % SIG)) * W];
    }
    A g[SIG];
#pragma unroll
    for (int q = 0; q < SIG; ++q) g[q] = (A)x[cc[q]];
#pragma unroll
    for (int q = 0; q < SIG; ++q) {
      const int j = q * 64 + lane;
      xb[(j / SIG) * (SIG + 1) + (j % SIG)] = g[q];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SIG; ++i) walk(i, v[i] * xb[lane * (SIG + 1) + i]);
  } else {
    // kChunk lane steps at a time: all column and value loads, then all x
    // gathers, then the flag walk, so that many loads are in flight
    constexpr int kChunk = 8;
    for (int i0 = 0; i0 < sigma; i0 += kChunk) {
      uint32_t c[kChunk];
      A v[kChunk], p[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int64_t e = base + (int64_t)(i0 + j) * 64;
        c[j] = i0 + j < sigma ? colp[e] : 0u;
        v[j] = i0 + j < sigma ? Val<VT>::get(val, e) : A(0);
      }
#pragma unroll
      for (int j = 0; j < kChunk; ++j) p[j] = i0 + j < sigma ? v[j] * (A)x[c[j]] : A(0);
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        if (i0 + j >= sigma) break;
        walk(i0 + j, p[j]);
      }
    }
  }
  const bool flagged = nfl > 0;
  if (!flagged) head = cur;
  // U[l] = head[l] + head[l+1] + ... up to and including the next flagged
  // lane (segmented suffix scan; a flagged lane closes its own chain)
  Ac U = head;
  int closed = flagged ? 1 : 0;
  for (int d = 1; d < 64; d <<= 1) {
    Ac ov;
    ov.s = __shfl_down(U.s, d, 64);
    if constexpr (COMP) ov.c = __shfl_down(U.c, d, 64);
    const int oc = __shfl_down(closed, d, 64);
    if (!closed && lane + d < 64) {
      U.merge(ov);
      closed |= oc;
    }
  }
  Ac cont;
  cont.s = __shfl_down(U.s, 1, 64);
  if constexpr (COMP) cont.c = __shfl_down(U.c, 1, 64);
  if (lane == 63) cont = Ac();
  const uint64_t mask = __ballot(flagged);
  const int last_lane = mask ? 63 - __builtin_clzll(mask) : -1;
  if (flagged) {
    Ac sc = cur;
    sc.merge(cont);
    const A s = sc.value();
    const int64_t r = row_of(yoff + nfl);
    if (lane != last_lane) {
      y[r] = alpha * s + scaled_y(beta, y, r);
    } else {
      last_part[t] = s;
      last_row[t] = (int32_t)r;
    }
  }
  if (lane == 0) {
    if (starts_here && mask) {
      y[row0] = alpha * U.value() + scaled_y(beta, y, row0);
    } else if (starts_here) {
      last_part[t] = U.value();
      last_row[t] = (int32_t)row0;
    } else {
      first_part[t] = U.value();
      if (!mask) last_row[t] = -1;
    }
  }
}

// One long segment or long tail row per wavefront: an open segment adds
// the last part of its tile, the leading parts of the tiles it spans and
// (if it runs into the tail) its tail entries, lanes striding and then a
// butterfly; a tail row is summed the same way.
template <class VT, bool COMP>
__device__ __forceinline__ void csr5_long_item(
    const Csr5Args& a, const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, const typename Val<VT>::acc* __restrict__ first_part,
    const typename Val<VT>::acc* __restrict__ last_part, const int32_t* __restrict__ last_row,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta, int64_t w) {
  using A = typename Val<VT>::acc;
  const int lane = threadIdx.x & 63;
  const int32_t code = a.long_fix[w];
  Acc<A, COMP> s;
  int64_t r, e0, e1;
  if (code >= 0) {
    const int64_t i = code;
    r = last_row[i];
    const uint32_t ch = (uint32_t)a.chain[i];
    const int64_t q1 = ch & 0x7FFFFFFFu;
    if (lane == 0) s.add(last_part[i]);
    for (int64_t q = i + 1 + lane; q < q1; q += 64) s.add(first_part[q]);
    e0 = (ch >> 31) ? a.tail_base : 0;
    e1 = (ch >> 31) ? (int64_t)a.rst[r + 1] : 0;
  } else {
    r = -(int64_t)code - 1;
    e0 = a.rst[r];
    e1 = a.rst[r + 1];
  }
  for (int64_t e = e0 + lane; e < e1; e += 64) s.add(Val<VT>::get(val, e) * (A)x[a.col[e]]);
  s = group_reduce(s, 64);
  if (lane == 0) y[r] = alpha * s.value() + scaled_y(beta, y, r);
}

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_csr5_fixup(
    Csr5Args a, const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, const typename Val<VT>::acc* __restrict__ first_part,
    const typename Val<VT>::acc* __restrict__ last_part, const int32_t* __restrict__ last_row,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta, int64_t fix_blocks) {
  using A = typename Val<VT>::acc;
  if ((int64_t)blockIdx.x >= fix_blocks) {  // the long items, a wavefront each
    const int64_t w = (blockIdx.x - fix_blocks) * (kWG / 64) + (threadIdx.x >> 6);
    if (w < a.nlong) csr5_long_item<VT, COMP>(a, val, x, y, first_part, last_part, last_row, alpha, beta, w);
    return;
  }
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  const int64_t ntail = a.m - a.first_tail_row;
  if (i < a.ntiles) {
    const int32_t r = last_row[i];
    if (r < 0) return;
    const uint32_t ch = (uint32_t)a.chain[i];
    const int64_t q1 = ch & 0x7FFFFFFFu;
    if ((ch >> 31) || q1 - i - 1 > kCsr5Serial) return;  // a long item
    Acc<A, COMP> s;
    s.add(last_part[i]);
    for (int64_t q = i + 1; q < q1; ++q) s.add(first_part[q]);
    y[r] = alpha * s.value() + scaled_y(beta, y, r);
  } else if (i < a.ntiles + ntail) {
    const int64_t r = a.first_tail_row + (i - a.ntiles);
    const int64_t e0 = a.rst[r], e1 = a.rst[r + 1];
    if (e0 == e1) return;  // empty rows are in the empty list
    if (e1 - e0 > kCsr5Serial) return;  // a long item
    Acc<A, COMP> s;
    for (int64_t e = e0; e < e1; ++e) s.add(Val<VT>::get(val, e) * (A)x[a.col[e]]);
    y[r] = alpha * s.value() + scaled_y(beta, y, r);
  } else if (i < a.ntiles + ntail + a.nempty) {
    const int64_t r = a.empty_rows[i - a.ntiles - ntail];
    y[r] = scaled_y(beta, y, r);
  }
}

template <class VT>
static int run_csr5(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevCSR5& d = A.csr5;
  Csr5Args a{d.tile_ptr, d.desc,   d.eptr,   d.eoff,    d.rst,        d.col,
             d.empty_rows, d.chain, d.long_fix, d.ntiles, d.nempty, A.m, d.nlong,
             d.sigma, d.dl, d.bit_y, d.bit_seg, d.tail_start, d.first_tail_row, d.tail_base};
  constexpr bool COMP = sizeof(VT) == 2;  // 16-bit values: TwoSum throughout
  AC* fp = (AC*)(scratch ? scratch : d.scratch);
  AC* lp = fp + d.ntiles;
  int32_t* lr = (int32_t*)(lp + d.ntiles);
  if (d.ntiles > 0) {
    if (d.sigma == 16) {
      hipLaunchKernelGGL((k_csr5_tiles<VT, 16, COMP>), dim3((unsigned)(sizeof(AC) < 8 ? xcd_grid(d.ntiles) : d.ntiles)), dim3(64), 0, s, a,
                         (const VT*)d.val, (const IO*)x, (IO*)y, fp, lp, lr, (AC)alpha, (AC)beta);
    } else if (d.sigma == 8) {
      hipLaunchKernelGGL((k_csr5_tiles<VT, 8, COMP>), dim3((unsigned)(sizeof(AC) < 8 ? xcd_grid(d.ntiles) : d.ntiles)), dim3(64), 0, s, a,
                         (const VT*)d.val, (const IO*)x, (IO*)y, fp, lp, lr, (AC)alpha, (AC)beta);
    } else {
      const unsigned g1 = (unsigned)((d.ntiles + kWG / 64 - 1) / (kWG / 64));
      hipLaunchKernelGGL((k_csr5_tiles<VT, 0, COMP>), dim3(g1), dim3(kWG), 0, s, a, (const VT*)d.val,
                         (const IO*)x, (IO*)y, fp, lp, lr, (AC)alpha, (AC)beta);
    }
    if (int rc = hip_check(hipGetLastError(), "k_csr5_tiles launch")) return rc;
  }
  // the fix-ups (a thread each) and the long items (a wavefront each) in one launch
  const int64_t nfix = d.ntiles + (A.m - d.first_tail_row) + d.nempty;
  const int64_t g2 = (nfix + kWG - 1) / kWG, g3 = (d.nlong + kWG / 64 - 1) / (kWG / 64);
  if (g2 + g3 > 0) {
    hipLaunchKernelGGL((k_csr5_fixup<VT, COMP>), dim3((unsigned)(g2 + g3)), dim3(kWG), 0, s, a,
                       (const VT*)d.val, (const IO*)x, (IO*)y, fp, lp, lr, (AC)alpha, (AC)beta, g2);
    return hip_check(hipGetLastError(), "k_csr5_fixup launch");
  }
  return RVVLA_OK;
}

int launch_csr5(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_csr5<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_csr5<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_csr5<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_csr5<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
