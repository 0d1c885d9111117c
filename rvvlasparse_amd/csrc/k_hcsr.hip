// k_hcsr.hip -- HCSR SpMV for gfx950.
//
// Semantics follow reference src/scalar/spmv_mtx_hcsr.cpp:52-76: every
// block of block-row br adds alpha * (block . x[bc*C ...]) to y[br*R ...],
// after y has been scaled by beta; here y = alpha * acc + beta * y, with y
// read and written once per row and only the blocks holding entries
// visited.  The image (hcsr_image.cpp) gives one wavefront per unit of a
// 64-row slice of a block-row, lane l owning the slice's row l:
//  * dense runs: slot row j of a run holds each lane's j-th entry of one
//    block (16-bit column inside the block, 0xFFFF for none; each slot row
//    carries the block's first column), so the column and value loads of
//    a slot row are whole 128- and 512-byte lines and the x gathers stay
//    inside the block's C columns; a unit's slot rows are consecutive;
//  * the tail, ordered by lane, 64 entries at a time (loads and gathers
//    four chunks ahead): a DPP segmented scan
//    over the chunk (segments = runs of one lane's entries) and the last
//    lane of each segment adds the segment's sum to its row's slot in a
//    64-entry LDS array of the wavefront -- at most one segment per row
//    in a chunk, so no two lanes touch one slot and no atomics are used.
// The order of every addition is fixed by the image, so results repeat
// bit for bit.  16-bit values accumulate with TwoSum compensation.  A
// slice cut into several units leaves one partial sum per row and unit,
// summed in unit order by k_hcsr_combine.  With 32-bit accumulation each
// XCD walks a contiguous run of units.  The common unit (at most 8 slot
// rows and one tail chunk) issues every load, then every gather, in one
// straight-line batch: for FP64 and the 16-bit forms a batch of exactly its
// slot rows (hc_fast, one instance per count), for FP32 a batch of 8 slot
// rows (measured faster for each).
#include <type_traits>

#include "kcommon.h"

namespace rvvla {

// One unit, any shape: its slot rows 8 at a time (offsets, columns and
// values all issued before the gathers; lanes past the slice's rows load
// nothing, their slots are empty), then its tail chunks.  Returns the
// lane's row sum.
template <class VT, bool COMP>
__device__ __forceinline__ Acc<typename Val<VT>::acc, COMP> hc_unit(
    int nrows, int s0, int s1, int te0, int te1, const int32_t* __restrict__ dxoff,
    const uint16_t* __restrict__ dcol, const VT* __restrict__ dval,
    const int32_t* __restrict__ tcol, const uint8_t* __restrict__ trow,
    const VT* __restrict__ tval, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::acc* ls, typename Val<VT>::acc* lc, int lane) {
  using A = typename Val<VT>::acc;
  using Ac = Acc<A, COMP>;
  Ac acc, acc1;  // even and odd slot rows
  if (lane < nrows) {
    for (int q0 = s0; q0 < s1; q0 += 8) {
      uint16_t c[8];
      A v[8];
      int32_t xo[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = min(q0 + k, s1 - 1);  // rows past s1 repeat the last one
        const int64_t slot = (int64_t)q * 64 + lane;
        xo[k] = dxoff[q];
        c[k] = dcol[slot];
        v[k] = Val<VT>::get(dval, slot);
      }
      A xv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        // an empty slot gathers the block's first x (in bounds) and adds 0
        xv[k] = (A)x[xo[k] + (c[k] != 0xFFFF ? c[k] : 0)];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        keep_live(xv[k]);
        const bool use = c[k] != 0xFFFF && q0 + k < s1;
        (k & 1 ? acc1 : acc).add(use ? v[k] * xv[k] : A(0));
      }
    }
  }
  acc.merge(acc1);
  if (te1 > te0) {
    ls[lane] = A(0);
    if constexpr (COMP) lc[lane] = A(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 4 chunks at a time: their loads and gathers first, then the scans
    const int elast = te0 + ((te1 - te0 - 1) & ~63);  // the last chunk's start
    for (int b0 = te0; b0 < te1; b0 += 4 * 64) {
      int r[4];
      A p[4], xv[4];
      bool valid[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = min(b0 + k * 64, elast) + lane;  // the tail is padded to whole chunks
        valid[k] = b0 + k * 64 + lane < te1;
        r[k] = trow[e];
        const int32_t c = tcol[e];
        const A v = Val<VT>::get(tval, e);
        xv[k] = (A)x[c];  // padding entries hold column 0
        p[k] = v;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        keep_live(xv[k]);
        p[k] = valid[k] ? p[k] * xv[k] : A(0);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (b0 + k * 64 >= te1) break;
        Ac S;
        S.s = p[k];
        const int key = valid[k] ? r[k] : 64 + lane;
        const int prev = __builtin_amdgcn_update_dpp(-1, key, 0x138, 0xF, 0xF, false);
        const uint64_t heads = __ballot(lane == 0 || key != prev);
        S = wave_seg_scan(S, heads, lane);
        const bool end = lane == 63 || ((heads >> (lane + 1)) & 1ull);
        if (valid[k] && end) {
          Ac t;
          t.s = ls[r[k]];
          if constexpr (COMP) t.c = lc[r[k]];
          t.merge(S);
          ls[r[k]] = t.s;
          if constexpr (COMP) lc[r[k]] = t.c;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    Ac t;
    t.s = ls[lane];
    if constexpr (COMP) t.c = lc[lane];
    acc.merge(t);
  }
  return acc;
}

// A wavefront per unit.  A unit that is a whole slice of at most 8 slot
// rows and at most one tail chunk (the common case for matrices whose
// blocks are dense enough) issues all its loads (slot rows, tail chunk, y)
// before any gather, and all its gathers before any sum (`fast`: the
// matrix has columns, so x[0] exists); other units run through hc_unit.
// The gathers are unconditional -- an empty slot gathers the first x of
// its block, a unit without slot rows reads the image's last slot row,
// which is always empty, with first column 0 -- and the products are
// selected, so no load sits under a branch.
// The common unit with exactly NS slot rows (NS <= 8) and at most one
// tail chunk: every load (slot rows, tail, y) in one straight-line batch,
// then every gather, then the sums (even and odd slot rows, as the general
// walk adds them).
template <class VT, bool COMP, int NS>
__device__ __forceinline__ void hc_fast(
    int32_t row0, int nrows, int s0, int te0, int te1, const int32_t* __restrict__ dxoff,
    const uint16_t* __restrict__ dcol, const VT* __restrict__ dval,
    const int32_t* __restrict__ tcol, const uint8_t* __restrict__ trow,
    const VT* __restrict__ tval, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc* ls, typename Val<VT>::acc* lc,
    int lane, typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  using Ac = Acc<A, COMP>;
  uint16_t c[NS > 0 ? NS : 1];
  A v[NS > 0 ? NS : 1], xv[NS > 0 ? NS : 1];
  int32_t xo[NS > 0 ? NS : 1];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int q = s0 + k;
    const int64_t slot = (int64_t)q * 64 + lane;
    xo[k] = dxoff[q];
    c[k] = dcol[slot];
    v[k] = Val<VT>::get(dval, slot);
  }
  const bool tail = te1 > te0;
  int tr = 0;
  int32_t tc = 0;  // x[0]
  A tv = A(0), yb = A(0), xt;
  if (tail) {  // the tail is padded to a whole chunk
    tr = trow[te0 + lane];
    tc = tcol[te0 + lane];
    tv = Val<VT>::get(tval, te0 + lane);
  }
  // lanes past the slice read its last row of y
  if (beta != A(0)) yb = (A)y[(int64_t)row0 + min(lane, nrows - 1)];
#pragma unroll
  for (int k = 0; k < NS; ++k) xv[k] = (A)x[xo[k] + (c[k] != 0xFFFF ? c[k] : 0)];
  xt = (A)x[tc];
  Ac acc[2];  // even and odd slot rows
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    keep_live(xv[k]);
    acc[k & 1].add(c[k] != 0xFFFF ? v[k] * xv[k] : A(0));
  }
  acc[0].merge(acc[1]);
  if (tail) {
    // one chunk: each row has at most one segment, whose last lane writes
    // the segment's sum to the row's LDS slot
    keep_live(xt);
    ls[lane] = A(0);
    if constexpr (COMP) lc[lane] = A(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool valid = te0 + lane < te1;
    Ac S;
    S.s = valid ? tv * xt : A(0);
    const int key = valid ? tr : 64 + lane;
    const int prev = __builtin_amdgcn_update_dpp(-1, key, 0x138, 0xF, 0xF, false);
    const uint64_t heads = __ballot(lane == 0 || key != prev);
    S = wave_seg_scan(S, heads, lane);
    if (valid && (lane == 63 || ((heads >> (lane + 1)) & 1ull))) {
      ls[tr] = S.s;
      if constexpr (COMP) lc[tr] = S.c;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    Ac t;
    t.s = ls[lane];
    if constexpr (COMP) t.c = lc[lane];
    acc[0].merge(t);
  }
  keep_live(yb);
  if (lane < nrows) y[(int64_t)row0 + lane] = alpha * acc[0].value() + beta * yb;
}

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_hcsr_slices(
    const int32_t* __restrict__ unit, int64_t nunits, const int32_t* __restrict__ dxoff,
    const uint16_t* __restrict__ dcol, const VT* __restrict__ dval,
    const int32_t* __restrict__ tcol, const uint8_t* __restrict__ trow,
    const VT* __restrict__ tval, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc* __restrict__ part,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta, int empty_row, bool fast) {
  using A = typename Val<VT>::acc;
  using Ac = Acc<A, COMP>;
  constexpr int NW = kWG / 64;
  __shared__ A ls[NW][64];
  __shared__ A lc[COMP ? NW : 1][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // 32-bit accumulation: each XCD walks a contiguous run of units
  const int wg = sizeof(A) < 8 ? xcd_run_index((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int64_t u = __builtin_amdgcn_readfirstlane((int)(wg * NW + wv));
  if (u >= nunits) return;
  const int32_t* U = unit + u * kHcUnitInts;
  const int32_t row0 = U[0], nrows = U[1], s0 = U[2], s1 = U[3], te0 = U[4], te1 = U[5],
                out = U[6];
  // the common unit: FP64 and the 16-bit forms run a batch of exactly its
  // slot rows (measured faster there), FP32 a batch of 8 slot rows, the
  // rows past the unit's repeating its last one
  constexpr bool EXACT = !std::is_same<VT, float>::value;
  if (EXACT && fast && s1 - s0 <= 8 && te1 - te0 <= 64 && out < 0) {
    switch (s1 - s0) {
#define RVVLA_HC_FAST(NS)                                                                        \
  case NS:                                                                                       \
    hc_fast<VT, COMP, NS>(row0, nrows, s0, te0, te1, dxoff, dcol, dval, tcol, trow, tval, x, y,   \
                          ls[wv], lc[COMP ? wv : 0], lane, alpha, beta);                         \
    return;
      RVVLA_HC_FAST(0) RVVLA_HC_FAST(1) RVVLA_HC_FAST(2) RVVLA_HC_FAST(3) RVVLA_HC_FAST(4)
      RVVLA_HC_FAST(5) RVVLA_HC_FAST(6) RVVLA_HC_FAST(7) RVVLA_HC_FAST(8)
#undef RVVLA_HC_FAST
      default: break;
    }
  }
  if (!EXACT && fast && s1 - s0 <= 8 && te1 - te0 <= 64 && out < 0) {
    uint16_t c[8];
    A v[8], xv[8];
    int32_t xo[8];
    const int last = s1 > s0 ? s1 - 1 : empty_row;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = min(s0 + k, last);  // rows past s1 repeat the last one
      const int64_t slot = (int64_t)q * 64 + lane;
      xo[k] = dxoff[q];
      c[k] = dcol[slot];
      v[k] = Val<VT>::get(dval, slot);
    }
    const bool tail = te1 > te0;
    int tr = 0;
    int32_t tc = 0;  // x[0]
    A tv = A(0), yb = A(0), xt;
    if (tail) {  // the tail is padded to a whole chunk
      tr = trow[te0 + lane];
      tc = tcol[te0 + lane];
      tv = Val<VT>::get(tval, te0 + lane);
    }
    // lanes past the slice read its last row of y
    if (beta != A(0)) yb = (A)y[(int64_t)row0 + min(lane, nrows - 1)];
#pragma unroll
    for (int k = 0; k < 8; ++k) xv[k] = (A)x[xo[k] + (c[k] != 0xFFFF ? c[k] : 0)];
    xt = (A)x[tc];
    Ac acc[2];  // even and odd slot rows
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      keep_live(xv[k]);
      const bool use = c[k] != 0xFFFF && s0 + k < s1;
      acc[k & 1].add(use ? v[k] * xv[k] : A(0));
    }
    acc[0].merge(acc[1]);
    if (tail) {
      // one chunk: each row has at most one segment, whose last lane
      // writes the segment's sum to the row's LDS slot
      keep_live(xt);
      ls[wv][lane] = A(0);
      if constexpr (COMP) lc[wv][lane] = A(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const bool valid = te0 + lane < te1;
      Ac S;
      S.s = valid ? tv * xt : A(0);
      const int key = valid ? tr : 64 + lane;
      const int prev = __builtin_amdgcn_update_dpp(-1, key, 0x138, 0xF, 0xF, false);
      const uint64_t heads = __ballot(lane == 0 || key != prev);
      S = wave_seg_scan(S, heads, lane);
      if (valid && (lane == 63 || ((heads >> (lane + 1)) & 1ull))) {
        ls[wv][tr] = S.s;
        if constexpr (COMP) lc[wv][tr] = S.c;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      Ac t;
      t.s = ls[wv][lane];
      if constexpr (COMP) t.c = lc[wv][lane];
      acc[0].merge(t);
    }
    keep_live(yb);
    if (lane < nrows) y[(int64_t)row0 + lane] = alpha * acc[0].value() + beta * yb;
    return;
  }
  const Ac acc = hc_unit<VT, COMP>(nrows, s0, s1, te0, te1, dxoff, dcol, dval, tcol, trow, tval,
                                   x, ls[wv], lc[COMP ? wv : 0], lane);
  if (lane < nrows) {
    const A sum = acc.value();
    if (out < 0) {
      const int64_t r = (int64_t)row0 + lane;
      y[r] = alpha * sum + scaled_y(beta, y, r);
    } else {
      part[(int64_t)out * 64 + lane] = sum;
    }
  }
}

// Slices cut into several units: y = alpha * (sum of the units' partial
// sums, in unit order) + beta * y, one wavefront per slice.
template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_hcsr_combine(
    const int32_t* __restrict__ split, int64_t nsplit, const typename Val<VT>::acc* __restrict__ part,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t q = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)));
  if (q >= nsplit) return;
  const int lane = threadIdx.x & 63;
  const int32_t row0 = split[q * 4], nrows = split[q * 4 + 1], p0 = split[q * 4 + 2],
                p1 = split[q * 4 + 3];
  Acc<A, COMP> s;
  for (int p = p0; p < p1; ++p) s.add(part[(int64_t)p * 64 + lane]);
  if (lane < nrows) {
    const int64_t r = (int64_t)row0 + lane;
    y[r] = alpha * s.value() + scaled_y(beta, y, r);
  }
}

template <class VT>
static int run_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  constexpr bool COMP = sizeof(VT) == 2;
  const DevHCSR& d = A.hcsr;
  AC* part = (AC*)(scratch ? scratch : d.part);
  if (d.nunits > 0) {
    const int64_t g0 = (d.nunits + kWG / 64 - 1) / (kWG / 64);
    const unsigned g = (unsigned)(sizeof(AC) < 8 ? xcd_grid(g0) : g0);
    hipLaunchKernelGGL((k_hcsr_slices<VT, COMP>), dim3(g), dim3(kWG), 0, s, d.unit, d.nunits,
                       d.dxoff, d.dcol, (const VT*)d.dval, d.tcol, d.trow,
                       (const VT*)d.tval, (const IO*)x, (IO*)y, part, (AC)alpha, (AC)beta,
                       (int)(d.ndrows - 1), A.n > 0);
    if (int rc = hip_check(hipGetLastError(), "k_hcsr_slices launch")) return rc;
  }
  if (d.nsplit > 0) {
    const unsigned g = (unsigned)((d.nsplit + kWG / 64 - 1) / (kWG / 64));
    hipLaunchKernelGGL((k_hcsr_combine<VT, COMP>), dim3(g), dim3(kWG), 0, s, d.split, d.nsplit,
                       (const AC*)part, (IO*)y, (AC)alpha, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_hcsr_combine launch")) return rc;
  }
  return RVVLA_OK;
}

int launch_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_hcsr<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_hcsr<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_hcsr<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_hcsr<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
