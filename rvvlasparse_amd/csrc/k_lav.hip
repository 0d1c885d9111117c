// k_lav.hip -- LAV SpMV for gfx950 (SIMD_Lanes = 64: one chunk per
// wavefront, one chunk row per lane).
//
// Semantics follow the reference scalar kernel (reference
// src/scalar/spmv_mtx_lav.cpp:23-79): first the CSR multiply of sparse_part
// with beta (here the adaptive CSR kernel), then every dense entry adds
// alpha * val * x[col] into its chunk row.  Because the library sorts rows
// once for all segments, a wavefront walks chunk c through every segment,
// keeps the lane sums in registers and adds alpha * sum into y once.  Slots
// past a step's mask were zeroed at upload, so all 64 lanes load every step
// (one contiguous access per step) without per-lane predicates.
#include "kcommon.h"

namespace rvvla {

template <class VT>
__global__ __launch_bounds__(kWG) void k_lav_dense(
    const int32_t* __restrict__ coff, const int64_t* __restrict__ sptr, int64_t nseg,
    int64_t nchunks, const uint32_t* __restrict__ out_order, const uint32_t* __restrict__ col,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, int64_t m, typename Val<VT>::acc alpha) {
  using A = typename Val<VT>::acc;
  const int64_t c = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (c >= nchunks) return;
  const int lane = threadIdx.x & 63;
  A a0 = 0, a1 = 0;
  for (int64_t g = 0; g < nseg; ++g) {
    const int64_t base = sptr[g];
    int64_t k = base + coff[g * (nchunks + 1) + c];
    const int64_t k1 = base + coff[g * (nchunks + 1) + c + 1];
    for (; k + 1 < k1; k += 2) {
      const int64_t e0 = k * 64 + lane, e1 = e0 + 64;
      const uint32_t c0 = col[e0], c1 = col[e1];
      a0 += Val<VT>::get(val, e0) * (A)x[c0];
      a1 += Val<VT>::get(val, e1) * (A)x[c1];
    }
    if (k < k1) {
      const int64_t e0 = k * 64 + lane;
      a0 += Val<VT>::get(val, e0) * (A)x[col[e0]];
    }
  }
  const int64_t rp = c * 64 + lane;
  if (rp < m) {
    const uint32_t r = out_order[rp];
    y[r] = (A)y[r] + alpha * (a0 + a1);
  }
}

template <class VT>
static int run_lav(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                   hipStream_t s) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  if (int rc = launch_csr(A, alpha, x, beta, y, s)) return rc;  // sparse_part, with beta
  const DevLAV& d = A.lav;
  if (d.nseg == 0 || d.nchunks == 0) return RVVLA_OK;
  const unsigned g = (unsigned)((d.nchunks + kWG / 64 - 1) / (kWG / 64));
  hipLaunchKernelGGL((k_lav_dense<VT>), dim3(g), dim3(kWG), 0, s, d.coff, d.sptr, d.nseg,
                     d.nchunks, d.out_order, d.col, (const VT*)d.val, (const IO*)x, (IO*)y, A.m,
                     (AC)alpha);
  return hip_check(hipGetLastError(), "k_lav_dense launch");
}

int launch_lav(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s) {
  switch (A.dtype) {
    case RVVLA_F64: return run_lav<double>(A, alpha, x, beta, y, s);
    case RVVLA_F32: return run_lav<float>(A, alpha, x, beta, y, s);
    case RVVLA_BF16: return run_lav<bf16_t>(A, alpha, x, beta, y, s);
    case RVVLA_F16: return run_lav<_Float16>(A, alpha, x, beta, y, s);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
