// k_sell.hip -- SELL-C-sigma SpMV for gfx950, C = 64.
//
// One wavefront per slice: lane k owns sorted row SBs[i] + k and walks the
// slice's Cl[i] slots; slot (j, k) is at Cs[i] + j*Bs[i] + k, so every load
// of values and columns is one contiguous 64-lane access.  The result goes
// straight to the original row through vPerm, fusing the reference's
// PREPARATION / GET_RESULTS permutations into the multiply as its vector
// path does (reference src/rvv/spmv_mtx_sell_c_sigma.cpp:60-85):
// y[perm[s]] = alpha * acc + beta * y[perm[s]].  When the permutation is
// the identity (sigma = 1) it is not read at all.  FP64 values and their
// columns are read with nontemporal loads.  Slices wider than kSellPiece
// steps (very uneven rows) are cut into pieces of kSellPiece steps run by
// their own wavefronts (k_sell_pieces); k_sell_combine adds a slice's
// pieces (a workgroup per slice, in an order fixed by the piece count)
// and writes its rows.
#include "kcommon.h"

namespace rvvla {

// One slice by one wavefront.  BOUND (the SELL image of a CSR matrix,
// rows in natural order): a lane adds only the entries of its row (its
// padding slots carry kSellPadBit in the column, whose product is dropped by
// a select), so that padding never multiplies x, as in the CSR product; the
// SELL format itself multiplies its zero padding like the reference's SELL
// kernels.
// A slice of exactly W steps (W <= 8): every column and value load, then
// every gather, with no predicate between them; the products are added as
// the general loop adds them (slot j to a[j mod 4] in whole groups of four,
// the rest to a0).
template <class VT, bool NT, bool BOUND, int W>
__device__ __forceinline__ void sell_fixed(int64_t idx, int height,
                                           const uint32_t* __restrict__ col,
                                           const VT* __restrict__ val,
                                           const typename Val<VT>::io* __restrict__ x,
                                           typename Val<VT>::acc (&a)[4]) {
  using A = typename Val<VT>::acc;
  uint32_t c[W];
  A v[W];
#pragma unroll
  for (int j = 0; j < W; ++j) {
    c[j] = ldu<NT>(col, idx + j * (int64_t)height);
    v[j] = ldv<NT>(val, idx + j * (int64_t)height);
  }
  A p[W];
#pragma unroll
  for (int j = 0; j < W; ++j) p[j] = v[j] * (A)x[BOUND ? c[j] & ~kSellPadBit : c[j]];
  constexpr int WB = W / 4 * 4;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const A q = !BOUND || !(c[j] & kSellPadBit) ? p[j] : A(0);
    if (j < WB) a[j % 4] += q;
    else a[0] += q;
  }
}

template <class VT, bool IDENT, bool NT, bool BOUND>
__device__ __forceinline__ void sell_slice(
    int64_t slice, int lane, const int32_t* __restrict__ cl, const int32_t* __restrict__ cs,
    const int32_t* __restrict__ bs, const int32_t* __restrict__ sbs,
    const uint32_t* __restrict__ col, const VT* __restrict__ val,
    const uint32_t* __restrict__ perm, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int height = bs[slice], width = cl[slice];
  if (width > kSellPiece) return;  // a wide slice: run as pieces
  const int64_t base = cs[slice];
  const int srow = sbs[slice];
  if (lane >= height) return;
  int64_t idx = base + lane;
  A a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  if (width <= 8) {  // uniform: one straight-line batch
    A a[4] = {A(0), A(0), A(0), A(0)};
    switch (width) {
      case 1: sell_fixed<VT, NT, BOUND, 1>(idx, height, col, val, x, a); break;
      case 2: sell_fixed<VT, NT, BOUND, 2>(idx, height, col, val, x, a); break;
      case 3: sell_fixed<VT, NT, BOUND, 3>(idx, height, col, val, x, a); break;
      case 4: sell_fixed<VT, NT, BOUND, 4>(idx, height, col, val, x, a); break;
      case 5: sell_fixed<VT, NT, BOUND, 5>(idx, height, col, val, x, a); break;
      case 6: sell_fixed<VT, NT, BOUND, 6>(idx, height, col, val, x, a); break;
      case 7: sell_fixed<VT, NT, BOUND, 7>(idx, height, col, val, x, a); break;
      case 8: sell_fixed<VT, NT, BOUND, 8>(idx, height, col, val, x, a); break;
      default: break;  // width 0: an empty slice
    }
    const uint32_t r = IDENT ? (uint32_t)(srow + lane) : perm[srow + lane];
    put_y(y, r, alpha * ((a[0] + a[1]) + (a[2] + a[3])) + scaled_y(beta, y, r));
    return;
  }
  int j = 0;
  for (; j + 3 < width; j += 4) {
    // all column and value loads first, then the four gathers
    const uint32_t c0 = ldu<NT>(col, idx), c1 = ldu<NT>(col, idx + height);
    const uint32_t c2 = ldu<NT>(col, idx + 2 * (int64_t)height);
    const uint32_t c3 = ldu<NT>(col, idx + 3 * (int64_t)height);
    const A v0 = ldv<NT>(val, idx), v1 = ldv<NT>(val, idx + height);
    const A v2 = ldv<NT>(val, idx + 2 * (int64_t)height);
    const A v3 = ldv<NT>(val, idx + 3 * (int64_t)height);
    // the gathers run for every slot (padding holds a valid column); BOUND
    // then drops a padding product (possibly 0 * Inf) by a select
    constexpr uint32_t K = BOUND ? ~kSellPadBit : ~0u;
    const A p0 = v0 * (A)x[c0 & K], p1 = v1 * (A)x[c1 & K], p2 = v2 * (A)x[c2 & K],
            p3 = v3 * (A)x[c3 & K];
    a0 += !BOUND || !(c0 & kSellPadBit) ? p0 : A(0);
    a1 += !BOUND || !(c1 & kSellPadBit) ? p1 : A(0);
    a2 += !BOUND || !(c2 & kSellPadBit) ? p2 : A(0);
    a3 += !BOUND || !(c3 & kSellPadBit) ? p3 : A(0);
    idx += 4 * (int64_t)height;
  }
  for (; j < width; ++j, idx += height) {
    const uint32_t c = ldu<NT>(col, idx);
    const A p = ldv<NT>(val, idx) * (A)x[BOUND ? c & ~kSellPadBit : c];
    a0 += !BOUND || !(c & kSellPadBit) ? p : A(0);
  }
  const uint32_t r = IDENT ? (uint32_t)(srow + lane) : perm[srow + lane];
  put_y(y, r, alpha * ((a0 + a1) + (a2 + a3)) + scaled_y(beta, y, r));
}

// CHUNK: a grid of at most 8 workgroups a CU, each walking its own
// contiguous run of 4-slice groups (its wavefronts side by side), the runs
// of consecutive workgroup indices laid out one XCD after another
// (workgroup b runs on XCD b mod 8): streams read by fewer, longer runs,
// measured faster for 32-bit and 16-bit values (2048^2 Laplacian FP32
// 36.8 -> 30.9 us).  Otherwise one slice per wavefront, slices in
// workgroup order (faster for FP64: 53 against 60 us).
template <class VT, bool IDENT, bool NT, bool BOUND, bool CHUNK>
__global__ __launch_bounds__(kWG) void k_sell64(
    const int32_t* __restrict__ cl, const int32_t* __restrict__ cs,
    const int32_t* __restrict__ bs, const int32_t* __restrict__ sbs,
    const uint32_t* __restrict__ col, const VT* __restrict__ val,
    const uint32_t* __restrict__ perm, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, int64_t nslices, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  const int lane = threadIdx.x & 63;
  // the slice index is uniform across the wavefront: keep it (and the slice
  // descriptor loads) scalar
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if constexpr (CHUNK) {
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;  // grid is a multiple of 8
    const int64_t bi = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t ng = (nslices + kWG / 64 - 1) / (kWG / 64), gpb = (ng + 8 * per - 1) / (8 * per);
    const int64_t g1 = min(ng, (bi + 1) * gpb);
    for (int64_t g = bi * gpb; g < g1; ++g) {
      const int64_t slice = g * (kWG / 64) + wave;
      if (slice >= nslices) break;
      sell_slice<VT, IDENT, NT, BOUND>(slice, lane, cl, cs, bs, sbs, col, val, perm, x, y, alpha,
                                       beta);
    }
  } else {
    const int64_t slice = (int64_t)blockIdx.x * (kWG / 64) + wave;
    if (slice >= nslices) return;
    sell_slice<VT, IDENT, NT, BOUND>(slice, lane, cl, cs, bs, sbs, col, val, perm, x, y, alpha,
                                     beta);
  }
}

// One piece (steps [j0, j1) of a wide slice) per wavefront: lane partials
// to part[piece * 64 + lane].  Wide slices hold the long rows: in the
// 16-bit value forms (COMP) every addition of a piece and of the combine is
// compensated (TwoSum) and a piece's partial is its compensated sum.
template <class VT, bool COMP, bool BOUND>
__global__ __launch_bounds__(kWG) void k_sell_pieces(
    const int32_t* __restrict__ pc_slice, const int32_t* __restrict__ pc_j0,
    const int32_t* __restrict__ pc_j1, int64_t npieces, const int32_t* __restrict__ cs,
    const int32_t* __restrict__ bs, const uint32_t* __restrict__ col, const VT* __restrict__ val,
    const typename Val<VT>::io* __restrict__ x, typename Val<VT>::acc* __restrict__ part,
    const int32_t* __restrict__ sbs, const int32_t* __restrict__ rst) {
  using A = typename Val<VT>::acc;
  const int64_t q = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (q >= npieces) return;
  const int lane = threadIdx.x & 63;
  const int sl = pc_slice[q], height = bs[sl];
  Acc<A, COMP> a0, a1, a2, a3;
  if (lane < height) {
    int j = pc_j0[q];
    int j1 = pc_j1[q];
    if constexpr (BOUND) {  // the row's own entries only
      const int r = sbs[sl] + lane;
      j1 = min(j1, rst[r + 1] - rst[r]);
    }
    int64_t idx = cs[sl] + (int64_t)j * height + lane;
    for (; j + 3 < j1; j += 4) {
      const uint32_t c0 = col[idx], c1 = col[idx + height];
      const uint32_t c2 = col[idx + 2 * (int64_t)height], c3 = col[idx + 3 * (int64_t)height];
      const A v0 = Val<VT>::get(val, idx), v1 = Val<VT>::get(val, idx + height);
      const A v2 = Val<VT>::get(val, idx + 2 * (int64_t)height);
      const A v3 = Val<VT>::get(val, idx + 3 * (int64_t)height);
      a0.add(v0 * (A)x[c0]);
      a1.add(v1 * (A)x[c1]);
      a2.add(v2 * (A)x[c2]);
      a3.add(v3 * (A)x[c3]);
      idx += 4 * (int64_t)height;
    }
    for (; j < j1; ++j, idx += height) a0.add(Val<VT>::get(val, idx) * (A)x[col[idx]]);
  }
  a0.merge(a1);
  a2.merge(a3);
  a0.merge(a2);
  part[q * 64 + lane] = a0.value();
}

// One wide slice per workgroup: the sum of its pieces' partials (all four
// wavefronts load pieces; a slice of the longest rows has hundreds), then
// the first wavefront writes the slice's rows.
template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_sell_combine(
    const int32_t* __restrict__ sp_slice, const int32_t* __restrict__ sp_first,
    const int32_t* __restrict__ bs, const int32_t* __restrict__ sbs,
    const uint32_t* __restrict__ perm, bool ident, const typename Val<VT>::acc* __restrict__ part,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  __shared__ A lds[COMP ? 2 * kWG : kWG];
  const int64_t w = blockIdx.x;
  const int sl = sp_slice[w];
  const A s = wg_piece_sum<A, COMP>(part, sp_first[w], sp_first[w + 1], lds);
  const int lane = threadIdx.x;
  if (lane >= bs[sl]) return;  // rows past the slice's height, and wavefronts 1..3
  const int k = sbs[sl] + lane;
  const uint32_t r = ident ? (uint32_t)k : perm[k];
  y[r] = alpha * s + scaled_y(beta, y, r);
}

template <class VT>
static int run_sell(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevSELL& d = A.sell;
  AC* part = (AC*)(scratch ? scratch : d.part);
  if (d.nslices == 0) return RVVLA_OK;
  constexpr bool CHUNK = sizeof(VT) < 8;
  int64_t grid = (d.nslices + kWG / 64 - 1) / (kWG / 64);
  if (CHUNK) grid = (std::min<int64_t>(grid, 8 * (int64_t)device_cu_count()) + 7) / 8 * 8;
  // nontemporal streaming pays for 8-byte values; narrower values run
  // faster with ordinary loads
  const bool nt = nt_enabled() && sizeof(VT) == 8;
  // the SELL image of a CSR matrix: rows bounded by their CSR lengths
  const bool bound = A.format == RVVLA_FMT_CSR;
#define RVVLA_SELL_LAUNCH(ID, NT_, B_)                                                          \
  hipLaunchKernelGGL((k_sell64<VT, ID, NT_, B_, CHUNK>), dim3((unsigned)grid), dim3(kWG), 0, s, \
                     d.cl, d.cs, d.bs, d.sbs, d.col, (const VT*)d.val, d.perm, (const IO*)x,    \
                     (IO*)y, d.nslices, (AC)alpha, (AC)beta)
  if (bound) {  // natural row order: the identity permutation
    if (nt) RVVLA_SELL_LAUNCH(true, true, true);
    else RVVLA_SELL_LAUNCH(true, false, true);
  } else if (d.perm_identity) {
    if (nt) RVVLA_SELL_LAUNCH(true, true, false);
    else RVVLA_SELL_LAUNCH(true, false, false);
  } else {
    if (nt) RVVLA_SELL_LAUNCH(false, true, false);
    else RVVLA_SELL_LAUNCH(false, false, false);
  }
#undef RVVLA_SELL_LAUNCH
  if (d.nsplit > 0) {
    if (int rc = hip_check(hipGetLastError(), "k_sell64 launch")) return rc;
    constexpr int W = kWG / 64;
    constexpr bool COMP = sizeof(VT) == 2;
    if (bound)
      hipLaunchKernelGGL((k_sell_pieces<VT, COMP, true>), dim3((unsigned)((d.npieces + W - 1) / W)),
                         dim3(kWG), 0, s, d.pc_slice, d.pc_j0, d.pc_j1, d.npieces, d.cs, d.bs,
                         d.col, (const VT*)d.val, (const IO*)x, part, d.sbs, A.csr.rst);
    else
      hipLaunchKernelGGL((k_sell_pieces<VT, COMP, false>), dim3((unsigned)((d.npieces + W - 1) / W)),
                         dim3(kWG), 0, s, d.pc_slice, d.pc_j0, d.pc_j1, d.npieces, d.cs, d.bs,
                         d.col, (const VT*)d.val, (const IO*)x, part, d.sbs, A.csr.rst);
    if (int rc = hip_check(hipGetLastError(), "k_sell_pieces launch")) return rc;
    hipLaunchKernelGGL((k_sell_combine<VT, COMP>), dim3((unsigned)d.nsplit), dim3(kWG), 0, s,
                       d.sp_slice, d.sp_first, d.bs, d.sbs, d.perm,
                       d.perm_identity, part, (IO*)y, (AC)alpha, (AC)beta);
  }
  return hip_check(hipGetLastError(), "k_sell64 launch");
}

int launch_sell(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_sell<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_sell<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_sell<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_sell<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
