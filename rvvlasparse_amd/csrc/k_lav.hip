// k_lav.hip -- LAV SpMV for gfx950 (SIMD_Lanes = 64: one chunk per
// wavefront, one chunk row per lane).
//
// Semantics follow the reference scalar kernel (reference
// src/scalar/spmv_mtx_lav.cpp:23-79): first the CSR multiply of sparse_part
// with beta (here the adaptive CSR kernel), then every dense entry adds
// alpha * val * x[col] into its chunk row.  When no sparse_part row has more
// than kLavFuseRow entries the two are fused: the lane that finishes a
// row's dense sum adds the row's sparse entries and writes
// y = alpha * (dense + sparse) + beta * y once (no CSR pass, no second
// pass over y).  Because the library sorts rows
// once for all segments, a wavefront walks chunk c through every segment,
// keeps the lane sums in registers and adds alpha * sum into y once.  All
// 64 lanes load every step (one contiguous access per step); the products
// of lanes past the step's lane count (its mask) are dropped by a select,
// so a padding slot never multiplies x (0 * Inf would be NaN).  Chunks
// longer than kLavPiece steps are cut into pieces (k_lav_pieces) whose
// partials k_lav_combine adds (a workgroup per chunk), so no wavefront
// walks a chunk of thousands of steps alone.
#include "kcommon.h"

namespace rvvla {

// The four gathers of a step group taken together before the lane selects
// (so that no select moves a gather under a branch of its own).
template <class A>
__device__ __forceinline__ void lav_keep(A& g0, A& g1, A& g2, A& g3) {
  __asm__ volatile("" : "+v"(g0), "+v"(g1), "+v"(g2), "+v"(g3));
}

// The sparse_part CSR arrays, for the fused form.
template <class VT>
struct LavSparse {
  const int32_t* rst;
  const int32_t* col;
  const VT* val;
};

// Fused form: the sparse_part sum of row r (at most kLavFuseRow entries).
template <class VT>
__device__ __forceinline__ typename Val<VT>::acc lav_sparse_row(
    const LavSparse<VT>& sp, int e0, int e1, const typename Val<VT>::io* __restrict__ x) {
  using A = typename Val<VT>::acc;
  A s = 0;
  // eight entries' loads, then their gathers, in flight together; added in order
  for (int e = e0; e < e1; e += 8) {
    int cc[8];
    A vv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      cc[j] = 0;
      vv[j] = A(0);
      if (e + j < e1) {
        cc[j] = sp.col[e + j];
        vv[j] = Val<VT>::get(sp.val, e + j);
      }
    }
    A p[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) p[j] = e + j < e1 ? vv[j] * (A)x[cc[j]] : A(0);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (e + j < e1) s += p[j];
  }
  return s;
}

template <class VT, bool FUSED, bool COMP>
__global__ __launch_bounds__(kWG) void k_lav_dense(
    const int32_t* __restrict__ coff, const int64_t* __restrict__ sptr, int64_t nseg,
    int64_t nchunks, const int32_t* __restrict__ whole, int64_t nwhole,
    const uint32_t* __restrict__ out_order, const uint32_t* __restrict__ col,
    const VT* __restrict__ val, const int32_t* __restrict__ cfull,
    const uint16_t* __restrict__ lanes, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, int64_t m, typename Val<VT>::acc alpha,
    LavSparse<VT> sp, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t w = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (w >= nwhole) return;
  const int64_t c = whole ? (int64_t)whole[w] : w;
  const int lane = threadIdx.x & 63;
  const int64_t rp = c * 64 + lane;
  // the lane's row (and, fused, its sparse_part bounds) before the dense walk
  uint32_t r = 0;
  int s0 = 0, s1 = 0;
  if (rp < m) {
    r = out_order[rp];
    if constexpr (FUSED) {
      s0 = sp.rst[r];
      s1 = sp.rst[r + 1];
    }
  }
  Acc<A, COMP> a0, a1, a2, a3;
  for (int64_t g = 0; g < nseg; ++g) {
    const int64_t base = sptr[g];
    int64_t k = base + coff[g * (nchunks + 1) + c];
    const int64_t k1 = base + coff[g * (nchunks + 1) + c + 1];
    // steps before kf hold all 64 lanes; from kf on, the products of lanes
    // past the step's lane count are dropped
    const int64_t kf = base + cfull[g * nchunks + c];
    for (; k + 3 < kf; k += 4) {
      const int64_t e0 = k * 64 + lane;
      const uint32_t c0 = col[e0], c1 = col[e0 + 64], c2 = col[e0 + 128], c3 = col[e0 + 192];
      const A v0 = Val<VT>::get(val, e0), v1 = Val<VT>::get(val, e0 + 64);
      const A v2 = Val<VT>::get(val, e0 + 128), v3 = Val<VT>::get(val, e0 + 192);
      a0.add(v0 * (A)x[c0]);
      a1.add(v1 * (A)x[c1]);
      a2.add(v2 * (A)x[c2]);
      a3.add(v3 * (A)x[c3]);
    }
    for (; k < kf; ++k) {
      const int64_t e0 = k * 64 + lane;
      a0.add(Val<VT>::get(val, e0) * (A)x[col[e0]]);
    }
    for (; k < k1; ++k) {
      const int64_t e0 = k * 64 + lane;
      A p = Val<VT>::get(val, e0) * (A)x[col[e0]];
      keep_live(p);
      a0.add(lane < lanes[k] ? p : A(0));
    }
  }
  if (rp < m) {
    a0.merge(a1);
    a2.merge(a3);
    a0.merge(a2);
    const A dense = a0.value();
    if constexpr (FUSED)
      put_y(y, r, alpha * (lav_sparse_row(sp, s0, s1, x) + dense) + scaled_y(beta, y, r));
    else
      y[r] = (A)y[r] + alpha * dense;
  }
}

// One piece (at most kLavPiece steps of one long chunk in one segment) per
// wavefront: its 64 lane sums go to part[piece * 64 + lane].  In the 16-bit
// value forms (COMP) the dense sums, pieces and combine add with TwoSum
// compensation (long chunks hold the long rows).
template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_lav_pieces(
    const int64_t* __restrict__ pc_k0, const int64_t* __restrict__ pc_k1, int64_t npieces,
    const uint32_t* __restrict__ col, const VT* __restrict__ val, const uint16_t* __restrict__ lanes,
    const typename Val<VT>::io* __restrict__ x, typename Val<VT>::acc* __restrict__ part) {
  using A = typename Val<VT>::acc;
  const int64_t q = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (q >= npieces) return;
  const int lane = threadIdx.x & 63;
  Acc<A, COMP> a0, a1, a2, a3;
  int64_t k = pc_k0[q];
  const int64_t k1 = pc_k1[q];
  for (; k + 3 < k1; k += 4) {
    const int64_t e0 = k * 64 + lane;
    const uint32_t c0 = col[e0], c1 = col[e0 + 64], c2 = col[e0 + 128], c3 = col[e0 + 192];
    const A v0 = Val<VT>::get(val, e0), v1 = Val<VT>::get(val, e0 + 64);
    const A v2 = Val<VT>::get(val, e0 + 128), v3 = Val<VT>::get(val, e0 + 192);
    const int n0 = lanes[k], n1 = lanes[k + 1], n2 = lanes[k + 2], n3 = lanes[k + 3];
    A g0 = (A)x[c0], g1 = (A)x[c1], g2 = (A)x[c2], g3 = (A)x[c3];
    lav_keep(g0, g1, g2, g3);
    a0.add(lane < n0 ? v0 * g0 : A(0));
    a1.add(lane < n1 ? v1 * g1 : A(0));
    a2.add(lane < n2 ? v2 * g2 : A(0));
    a3.add(lane < n3 ? v3 * g3 : A(0));
  }
  for (; k < k1; ++k) {
    const int64_t e0 = k * 64 + lane;
    A p = Val<VT>::get(val, e0) * (A)x[col[e0]];
    keep_live(p);
    a0.add(lane < lanes[k] ? p : A(0));
  }
  a0.merge(a1);
  a2.merge(a3);
  a0.merge(a2);
  part[q * 64 + lane] = a0.value();
}

// One split chunk per workgroup: the sum of its pieces' partials (all four
// wavefronts load pieces), then the first wavefront adds alpha * sum into
// the chunk's rows.
template <class VT, bool FUSED, bool COMP>
__global__ __launch_bounds__(kWG) void k_lav_combine(
    const int32_t* __restrict__ sp_chunk, const int32_t* __restrict__ sp_first,
    const typename Val<VT>::acc* __restrict__ part, const uint32_t* __restrict__ out_order,
    typename Val<VT>::io* __restrict__ y, int64_t m, typename Val<VT>::acc alpha,
    LavSparse<VT> sp, const typename Val<VT>::io* __restrict__ x, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  __shared__ A lds[COMP ? 2 * kWG : kWG];
  const int64_t w = blockIdx.x;
  const A s = wg_piece_sum<A, COMP>(part, sp_first[w], sp_first[w + 1], lds);
  if (threadIdx.x >= 64) return;
  const int64_t rp = (int64_t)sp_chunk[w] * 64 + threadIdx.x;
  if (rp < m) {
    const uint32_t r = out_order[rp];
    if constexpr (FUSED)
      y[r] = alpha * (lav_sparse_row(sp, sp.rst[r], sp.rst[r + 1], x) + s) + scaled_y(beta, y, r);
    else
      y[r] = (A)y[r] + alpha * s;
  }
}

template <class VT>
static int run_lav(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                   hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevLAV& d = A.lav;
  AC* part = (AC*)(scratch ? scratch : d.part);
  // sparse_part with beta first, unless each row's sparse entries are added
  // by the dense kernels (fused)
  if (!d.fused)
    if (int rc = launch_csr(A, alpha, x, beta, y, s)) return rc;
  if (d.nseg == 0 || d.nchunks == 0) return RVVLA_OK;
  const LavSparse<VT> sp{A.csr.rst, A.csr.col, (const VT*)A.csr.val};
  constexpr int W = kWG / 64;
  constexpr bool COMP = sizeof(VT) == 2;
  if (d.nwhole > 0) {
    const unsigned g = (unsigned)((d.nwhole + W - 1) / W);
#define RVVLA_LAV_DENSE(F)                                                                      \
  hipLaunchKernelGGL((k_lav_dense<VT, F, COMP>), dim3(g), dim3(kWG), 0, s, d.coff, d.sptr,     \
                     d.nseg, d.nchunks, d.whole, d.nwhole, d.out_order, d.col,                \
                     (const VT*)d.val, d.cfull, d.mask, (const IO*)x, (IO*)y, A.m, (AC)alpha, sp, \
                     (AC)beta)
    if (d.fused) RVVLA_LAV_DENSE(true);
    else RVVLA_LAV_DENSE(false);
#undef RVVLA_LAV_DENSE
    if (int rc = hip_check(hipGetLastError(), "k_lav_dense launch")) return rc;
  }
  if (d.nsplit > 0) {
    hipLaunchKernelGGL((k_lav_pieces<VT, COMP>), dim3((unsigned)((d.npieces + W - 1) / W)), dim3(kWG),
                       0, s, d.pc_k0, d.pc_k1, d.npieces, d.col, (const VT*)d.val, d.mask,
                       (const IO*)x, part);
    if (int rc = hip_check(hipGetLastError(), "k_lav_pieces launch")) return rc;
    if (d.fused)
      hipLaunchKernelGGL((k_lav_combine<VT, true, COMP>), dim3((unsigned)d.nsplit), dim3(kWG), 0, s,
                         d.sp_chunk, d.sp_first, part, d.out_order, (IO*)y, A.m,
                         (AC)alpha, sp, (const IO*)x, (AC)beta);
    else
      hipLaunchKernelGGL((k_lav_combine<VT, false, COMP>), dim3((unsigned)d.nsplit), dim3(kWG), 0, s,
                         d.sp_chunk, d.sp_first, part, d.out_order, (IO*)y, A.m,
                         (AC)alpha, sp, (const IO*)x, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_lav_combine launch")) return rc;
  }
  return RVVLA_OK;
}

int launch_lav(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_lav<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_lav<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_lav<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_lav<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
