// k_csr.hip -- CSR SpMV / SpMM for gfx950.
//
// Semantics: y_i = alpha * sum_j Val_j * x[Col_j] + beta * y_i, as the
// reference scalar CSR (reference src/scalar/spmv_mtx_crs.cpp:37-43), except
// that y is not read when beta == 0.
//
// One workgroup (256 threads = 4 wavefronts) per row block, built once at
// upload:
//  * stream block: up to 256 consecutive rows with <= 2048 entries in all.
//    The entries are read fully coalesced, multiplied by x and staged in
//    LDS; then a row of at most kShortSum entries is summed out of LDS by
//    its own thread in entry order, a longer one by a whole wavefront
//    (strided lanes, then a butterfly).
//  * long block: one row with more than 1024 entries; all 256 threads
//    stride over it, then wave and workgroup reductions.  In the 16-bit
//    value forms the partial sums are compensated (TwoSum) there.
//
// Blocks are (first, end) row pairs, long-row blocks first (longest first)
// so that no long row starts at the end of the launch.
//
// Column-sliced image (csr_xslice.cpp): the same kernel runs over the
// compressed (row, slice) rows, each block on the XCD of its slice so that
// each XCD gathers from its own slice of x (entries and columns are then streamed
// with nontemporal loads, to leave L2 to x), and writes one partial per pair
// (alpha = 1, beta = 0); k_xs_combine then adds each row's partials in
// slice order and applies alpha and beta.
#include "kcommon.h"

namespace rvvla {

constexpr int kShortSum = 16;

template <class VT, bool COMP, bool NT>
__global__ __launch_bounds__(kWG) void k_csr_adaptive(
    const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const int32_t* __restrict__ blocks,
    const typename Val<VT>::io* __restrict__ x, typename Val<VT>::io* __restrict__ y,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  __shared__ A prod[kStreamCap];
  __shared__ A wsum[2 * (kWG / 64)];
  const int tid = threadIdx.x;
  const int r0 = blocks[2 * blockIdx.x], r1 = blocks[2 * blockIdx.x + 1];
  if (r0 == r1) return;  // padding block of a sliced image (uniform)
  const int e0 = rst[r0], e1 = rst[r1];
  const int nnz = e1 - e0, nrows = r1 - r0;

  if (nrows == 1 && nnz > kLongRow) {
    Acc<A, COMP> a;
    int e = e0 + tid;
    for (; e + 3 * kWG < e1; e += 4 * kWG) {
      const int c0 = ldi<NT>(col, e), c1 = ldi<NT>(col, e + kWG);
      const int c2 = ldi<NT>(col, e + 2 * kWG), c3 = ldi<NT>(col, e + 3 * kWG);
      const A v0 = ldv<NT>(val, e), v1 = ldv<NT>(val, e + kWG);
      const A v2 = ldv<NT>(val, e + 2 * kWG), v3 = ldv<NT>(val, e + 3 * kWG);
      a.add(v0 * (A)x[c0]);
      a.add(v1 * (A)x[c1]);
      a.add(v2 * (A)x[c2]);
      a.add(v3 * (A)x[c3]);
    }
    for (; e < e1; e += kWG) a.add(ldv<NT>(val, e) * (A)x[ldi<NT>(col, e)]);
    a = group_reduce(a, 64);
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
      wsum[w] = a.s;
      wsum[kWG / 64 + w] = a.c;
    }
    __syncthreads();
    if (tid == 0) {
      Acc<A, COMP> t;
      t.s = wsum[0];
      t.c = wsum[kWG / 64];
      for (int i = 1; i < kWG / 64; ++i) {
        Acc<A, COMP> o;
        o.s = wsum[i];
        o.c = wsum[kWG / 64 + i];
        t.merge(o);
      }
      y[r0] = alpha * t.value() + scaled_y(beta, y, r0);
    }
    return;
  }

  // stream block: row bounds of thread tid's row (fetched before the
  // barrier), then coalesced products into LDS
  const int t_row = r0 + tid;
  int lo = 0, hi = 0;
  if (tid < nrows) {
    lo = rst[t_row] - e0;
    hi = rst[t_row + 1] - e0;
  }
  {
    int k = tid;
    for (; k + 3 * kWG < nnz; k += 4 * kWG) {
      const int c0 = ldi<NT>(col, e0 + k), c1 = ldi<NT>(col, e0 + k + kWG);
      const int c2 = ldi<NT>(col, e0 + k + 2 * kWG), c3 = ldi<NT>(col, e0 + k + 3 * kWG);
      const A v0 = ldv<NT>(val, e0 + k), v1 = ldv<NT>(val, e0 + k + kWG);
      const A v2 = ldv<NT>(val, e0 + k + 2 * kWG), v3 = ldv<NT>(val, e0 + k + 3 * kWG);
      prod[k] = v0 * (A)x[c0];
      prod[k + kWG] = v1 * (A)x[c1];
      prod[k + 2 * kWG] = v2 * (A)x[c2];
      prod[k + 3 * kWG] = v3 * (A)x[c3];
    }
    for (; k < nnz; k += kWG) prod[k] = ldv<NT>(val, e0 + k) * (A)x[ldi<NT>(col, e0 + k)];
  }
  __syncthreads();
  // rows of at most kShortSum products: their own thread, in entry order;
  // longer rows: the whole wavefront of the row's thread, one row at a time
  const bool is_long = tid < nrows && hi - lo > kShortSum;
  if (tid < nrows && !is_long) {
    A s = 0;
    for (int k = lo; k < hi; ++k) s += prod[k];
    y[t_row] = alpha * s + scaled_y(beta, y, t_row);
  }
  const int lane = tid & 63;
  uint64_t lm = __ballot(is_long);
  while (lm) {
    const int l = __ffsll((unsigned long long)lm) - 1;
    lm &= lm - 1;
    const int rlo = __shfl(lo, l, 64), rhi = __shfl(hi, l, 64);
    Acc<A, false> a;
    for (int k = rlo + lane; k < rhi; k += 64) a.add(prod[k]);
    a = group_reduce(a, 64);
    if (lane == 0) {
      const int r = r0 + (tid - lane) + l;
      y[r] = alpha * a.value() + scaled_y(beta, y, r);
    }
  }
}

// y[r] = alpha * (sum over slices p in order of part[pair(r, p)]) + beta*y[r].
// A wavefront covers kCombineGroups consecutive groups of 64 rows, one row
// of each per lane; the pair of row r (group q) in slice p is base[p] +
// pref[p*nq + q] + (lower lanes with bit p set).  All partial loads of all
// groups are issued before the (in-order) sums.
constexpr int kCombineGroups = 1;
template <class IO, class A>
__global__ __launch_bounds__(kWG) void k_xs_combine(
    const uint16_t* __restrict__ mask, const int32_t* __restrict__ base,
    const int32_t* __restrict__ pref, const IO* __restrict__ part, IO* __restrict__ y,
    int64_t m, int64_t nq, int P, A alpha, A beta) {
  constexpr int G = kCombineGroups;
  const int lane = threadIdx.x & 63;
  const int64_t q0 =
      ((int64_t)blockIdx.x * (kWG / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * G;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned mk[G];
  A v[G][16];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t r = (q0 + g) * 64 + lane;
    mk[g] = r < m ? mask[r] : 0u;
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const bool has = p < P && ((mk[g] >> p) & 1u);
      const uint64_t bal = __ballot(has);
      v[g][p] = A(0);
      if (has) v[g][p] = (A)part[base[p] + pref[(int64_t)p * nq + q0 + g] + __popcll(bal & below)];
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t r = (q0 + g) * 64 + lane;
    A s = A(0);
#pragma unroll
    for (int p = 0; p < 16; ++p)
      if ((mk[g] >> p) & 1u) s += v[g][p];
    if (r < m) y[r] = alpha * s + scaled_y(beta, y, r);
  }
}

// LDS-sliced image (csr_xlds.cpp).  One workgroup per CU runs its units
// (slice p, pairs g0 .. g1) in turn: x[p W, p W + W) into LDS (when the
// slice changes), then tiles of kXlTile pairs claimed by the wavefronts in
// turn.  A wavefront streams its tile's entries 64 at a time (kXlAhead
// chunks of loads in flight), each lane one entry; bit 15 of the column
// marks the first entry of a pair, so a ballot and a count of the marks
// below each lane give every entry its pair.  Each lane keeps the sum of
// its entries of its current pair in a register and adds it to the pair's
// LDS accumulator when its pair changes (lanes in order, so the order of
// the additions is fixed by the data); at the end of the tile the
// accumulators go to the pairs' slots.
constexpr int kXlAhead = 8;
static_assert(xl_wg_per_cu(8) >= 1 && xl_wg_per_cu(4) >= 1, "LDS of k_xl_slices");
template <class VT>
__global__ __launch_bounds__(kXlWG) void k_xl_slices(
    const uint16_t* __restrict__ col, const VT* __restrict__ val, const int32_t* __restrict__ pst,
    const int32_t* __restrict__ wpos, const int32_t* __restrict__ units,
    const int32_t* __restrict__ wg, const typename Val<VT>::io* __restrict__ x, int64_t n,
    int64_t W, typename Val<VT>::acc* __restrict__ part) {
  using A = typename Val<VT>::acc;
  using IO = typename Val<VT>::io;
  constexpr int XW = kXlSliceCols;
  constexpr int XP = XW / kXlWG;
  static_assert(XW % kXlWG == 0, "x slice per thread");
  __shared__ IO xs[XW];
  // pair accumulators in double for every value type: LDS FP32 atomic adds
  // kept the wavefronts waiting nine times longer than FP64 ones on MI355X
  __shared__ double acc[(kXlWG / 64) * kXlTile];
  __shared__ int next_tile;
  const int tid = threadIdx.x, lane = tid & 63;
  double* pacc = acc + (tid >> 6) * kXlTile;
  const int u0 = wg[blockIdx.x], u1 = wg[blockIdx.x + 1];
  int64_t loaded = -1;
  uint32_t ca[kXlAhead];
  A va[kXlAhead];
  for (int u = u0; u < u1; ++u) {
    const int64_t p = units[3 * u];
    const int64_t g0 = units[3 * u + 1], g1 = units[3 * u + 2];
    __syncthreads();  // every wavefront is done with the previous unit
    if (p != loaded) {
      const int64_t c0 = p * W;
      const int wl = (int)(n - c0 < W ? n - c0 : W);
      IO tmp[XP];
#pragma unroll
      for (int j = 0; j < XP; ++j) {
        const int i = tid + j * kXlWG;
        tmp[j] = i < wl ? x[c0 + i] : IO(0);
      }
#pragma unroll
      for (int j = 0; j < XP; ++j) xs[tid + j * kXlWG] = tmp[j];
      loaded = p;
    }
    if (tid == 0) next_tile = kXlWG / 64;  // tiles 0 .. waves-1 go one per wavefront
    __syncthreads();
    const int64_t ntiles = (g1 - g0 + kXlTile - 1) / kXlTile;
    int64_t t = tid >> 6;
    while (t < ntiles) {
      int nt = 0;  // the next tile, claimed now
      if (lane == 0) nt = atomicAdd(&next_tile, 1);
      const int64_t gt = g0 + t * kXlTile;
      const int np = (int)(gt + kXlTile < g1 ? kXlTile : g1 - gt);
      const int e_beg = pst[gt], e_end = pst[gt + np];
      for (int i = lane; i < kXlTile; i += 64) pacc[i] = 0.0;
      int cur = -1;   // tile-local pair of the lane's running sum
      A run = A(0);
      int open = -1;  // tile-local pair open before the chunk
      for (int cb = e_beg; cb < e_end; cb += 64 * kXlAhead) {
        // loads without branches (an index past the tile reads its last
        // entry and is masked), so no wait for all loads lands among them
#pragma unroll
        for (int k = 0; k < kXlAhead; ++k) {
          const int e = cb + k * 64 + lane < e_end ? cb + k * 64 + lane : e_end - 1;
          ca[k] = (uint32_t)col[e];
          va[k] = Val<VT>::get(val, e);
        }
        A prod[kXlAhead];  // the group's x gathers issued together
#pragma unroll
        for (int k = 0; k < kXlAhead; ++k) prod[k] = va[k] * (A)xs[ca[k] & 0x7FFFu];
#pragma unroll
        for (int k = 0; k < kXlAhead; ++k) {
          const bool in = cb + k * 64 + lane < e_end;
          const uint64_t H = __ballot(in && (ca[k] & 0x8000u));
          const int below = (int)__builtin_amdgcn_mbcnt_hi(
              (uint32_t)(H >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)H, 0u));
          const int lp = open + below + (int)((H >> lane) & 1ull);
          if (in) {
            if (lp != cur) {
              if (cur >= 0) atomicAdd(&pacc[cur], (double)run);
              run = A(0);
              cur = lp;
            }
            run += prod[k];
          }
          open += __popcll(H);
        }
      }
      if (cur >= 0) atomicAdd(&pacc[cur], (double)run);
      for (int i = lane; i < np; i += 64) part[wpos[gt + i]] = (A)pacc[i];
      t = __shfl(nt, 0, 64);
    }
  }
}

// Combine block b of the LDS-sliced image: rows r0 .. r1-1 (at most
// xl_comb_rows, one per thread) with slots base .. end-1.  Each thread
// loads its (at most 8) partials together and puts each at its row-major
// place in LDS (lpos); then each row's partials are added in slice order
// and alpha and beta applied.  A block of one row with more than
// xl_comb_pairs partials (then already in slice order) is summed by the
// whole workgroup.
template <class A, class IO>
__global__ __launch_bounds__(xl_comb_rows(sizeof(A))) void k_xl_combine(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ cblk,
                                                      const uint16_t* __restrict__ lpos,
                                                      const A* __restrict__ part,
                                                      IO* __restrict__ y, A alpha, A beta) {
  constexpr int kXlCombWG = xl_comb_rows(sizeof(A));
  constexpr int kXlCombPairs = xl_comb_pairs(sizeof(A));
  constexpr int PT = kXlCombPairs / kXlCombWG;
  __shared__ A prod[kXlCombPairs];
  __shared__ A wsum[kXlCombWG / 64];
  const int tid = threadIdx.x;
  const int r0 = cblk[4 * blockIdx.x], r1 = cblk[4 * blockIdx.x + 1];
  const int base = cblk[4 * blockIdx.x + 2], cnt = cblk[4 * blockIdx.x + 3];
  if (cnt > kXlCombPairs) {
    Acc<A, false> a;
    for (int q = tid; q < cnt; q += kXlCombWG) a.add(part[base + q]);
    a = group_reduce(a, 64);
    if ((tid & 63) == 0) wsum[tid >> 6] = a.s;
    __syncthreads();
    if (tid == 0) {
      A t = wsum[0];
#pragma unroll
      for (int w = 1; w < kXlCombWG / 64; ++w) t += wsum[w];
      y[r0] = alpha * t + scaled_y(beta, y, r0);
    }
    return;
  }
  const int r = r0 + tid;  // this thread's row: its bounds fetched with the partials
  int k0 = 0, k1 = 0;
  if (r < r1) {
    k0 = rowptr[r] - base;
    k1 = rowptr[r + 1] - base;
  }
  A v[PT];
  int at[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const int q = tid + j * kXlCombWG;
    const int qq = q < cnt ? q : (cnt > 0 ? cnt - 1 : 0);  // in range, masked below
    v[j] = cnt > 0 ? part[base + qq] : A(0);
    at[j] = cnt > 0 ? (int)lpos[base + qq] : 0;
  }
#pragma unroll
  for (int j = 0; j < PT; ++j)
    if (tid + j * kXlCombWG < cnt) prod[at[j]] = v[j];
  __syncthreads();
  if (r < r1) {
    A s = A(0);
    for (int k = k0; k < k1; ++k) s += prod[k];
    y[r] = alpha * s + scaled_y(beta, y, r);
  }
}

// SpMM: Y = alpha*A*X + beta*Y, X n x k, Y m x k (row-major).  One
// wavefront per row; lanes cover the k right-hand sides (k <= 64 per pass),
// so X rows are read as contiguous k-vectors.
template <class VT>
__global__ __launch_bounds__(kWG) void k_csr_spmm(
    const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ X, int64_t ldx,
    typename Val<VT>::io* __restrict__ Y, int64_t ldy, int64_t m, int k,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t row = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  const int e0 = rst[row], e1 = rst[row + 1];
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int c = c0 + lane;
    if (c >= k) break;
    A s = 0;
    for (int e = e0; e < e1; ++e) s += Val<VT>::get(val, e) * (A)X[(int64_t)col[e] * ldx + c];
    A* yp = &Y[row * ldy + c];
    *yp = alpha * s + scaled_y(beta, yp, 0);
  }
}

template <class VT>
static int run_csr(const DevCSR& d, int64_t m, double alpha, const void* x, double beta,
                   void* y, hipStream_t s) {
  const bool xs_nt = nt_enabled();
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const bool comp = sizeof(VT) == 2;
  if (d.xl.S > 0) {
    const DevXLds& xl = d.xl;
    if (m == 0) return RVVLA_OK;
    if (xl.nunits > 0) {
      hipLaunchKernelGGL((k_xl_slices<VT>), dim3((unsigned)xl.nwg), dim3(kXlWG), 0, s, xl.col,
                         (const VT*)xl.val, xl.pst, xl.wpos, xl.units, xl.wg, (const IO*)x, xl.n,
                         xl.W, (AC*)xl.part);
      if (int rc = hip_check(hipGetLastError(), "k_xl_slices launch")) return rc;
    }
    hipLaunchKernelGGL((k_xl_combine<AC, IO>), dim3((unsigned)xl.nblocks),
                       dim3(xl_comb_rows(sizeof(AC))), 0, s,
                       xl.rowptr, xl.cblk, xl.lpos, (const AC*)xl.part, (IO*)y, (AC)alpha,
                       (AC)beta);
    return hip_check(hipGetLastError(), "k_xl_combine launch");
  }
  if (d.xs.P > 0) {
    const DevXSlice& xs = d.xs;
    if (m == 0) return RVVLA_OK;
    if (xs.nblk > 0) {
      if (comp)
        hipLaunchKernelGGL((k_csr_adaptive<VT, true, true>), dim3((unsigned)xs.nblk), dim3(kWG), 0,
                           s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           (IO*)xs.part, (AC)1, (AC)0);
      else if (xs_nt)
        hipLaunchKernelGGL((k_csr_adaptive<VT, false, true>), dim3((unsigned)xs.nblk), dim3(kWG),
                           0, s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           (IO*)xs.part, (AC)1, (AC)0);
      else
        hipLaunchKernelGGL((k_csr_adaptive<VT, false, false>), dim3((unsigned)xs.nblk), dim3(kWG),
                           0, s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           (IO*)xs.part, (AC)1, (AC)0);
      if (int rc = hip_check(hipGetLastError(), "k_csr_adaptive (sliced) launch")) return rc;
    }
    const int64_t per_wg = (int64_t)(kWG / 64) * kCombineGroups * 64;
    const unsigned g = (unsigned)((m + per_wg - 1) / per_wg);
    hipLaunchKernelGGL((k_xs_combine<IO, AC>), dim3(g), dim3(kWG), 0, s, xs.mask, xs.base, xs.pref,
                       (const IO*)xs.part, (IO*)y, m, xs.nq, xs.P, (AC)alpha, (AC)beta);
    return hip_check(hipGetLastError(), "k_xs_combine launch");
  }
  if (d.nblocks == 0) return RVVLA_OK;
  if (comp)
    hipLaunchKernelGGL((k_csr_adaptive<VT, true, false>), dim3((unsigned)d.nblocks), dim3(kWG), 0,
                       s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  else if (xs_nt && sizeof(VT) == 8)  // FP64 entries streamed past L2
    hipLaunchKernelGGL((k_csr_adaptive<VT, false, true>), dim3((unsigned)d.nblocks), dim3(kWG),
                       0, s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  else
    hipLaunchKernelGGL((k_csr_adaptive<VT, false, false>), dim3((unsigned)d.nblocks), dim3(kWG),
                       0, s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  return hip_check(hipGetLastError(), "k_csr_adaptive launch");
}

int launch_csr_image(const DevCSR& d, int64_t m, int dtype, double alpha, const void* x,
                     double beta, void* y, hipStream_t s) {
  switch (dtype) {
    case RVVLA_F64: return run_csr<double>(d, m, alpha, x, beta, y, s);
    case RVVLA_F32: return run_csr<float>(d, m, alpha, x, beta, y, s);
    case RVVLA_BF16: return run_csr<bf16_t>(d, m, alpha, x, beta, y, s);
    case RVVLA_F16: return run_csr<_Float16>(d, m, alpha, x, beta, y, s);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

int launch_csr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s) {
  return launch_csr_image(A.csr, A.m, A.dtype, alpha, x, beta, y, s);
}

template <class VT>
static int run_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  if (A.m == 0 || k == 0) return RVVLA_OK;
  const unsigned grid = (unsigned)((A.m + kWG / 64 - 1) / (kWG / 64));
  hipLaunchKernelGGL((k_csr_spmm<VT>), dim3(grid), dim3(kWG), 0, s, A.csr.rst, A.csr.col,
                     (const VT*)A.csr.val, (const IO*)X, ldx, (IO*)Y, ldy, A.m, k, (AC)alpha,
                     (AC)beta);
  return hip_check(hipGetLastError(), "k_csr_spmm launch");
}

int launch_csr_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s) {
  switch (A.dtype) {
    case RVVLA_F64: return run_spmm<double>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_F32: return run_spmm<float>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_BF16: return run_spmm<bf16_t>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_F16: return run_spmm<_Float16>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
