// k_hcsr.hip -- HCSR SpMV for gfx950.
//
// Semantics follow reference src/scalar/spmv_mtx_hcsr.cpp:52-76: every
// block of block-row br adds alpha * (block . x[bc*C ...]) to y[br*R ...],
// after y has been scaled by beta.  Here one workgroup owns a block-row:
// its R partial sums live in LDS, only the non-empty blocks are visited
// (the reference walks all b_n of them), and y is read and written once
// (a block-row with more than kHcsrFlat blocks is walked entry by entry,
// each entry's local row and block column precomputed at upload):
// y = alpha * acc + beta * y.
//  * CSR-kind block: thread per local row, local row_ptr from rp32,
//    16-bit columns relative to the block's first column.
//  * COO-kind block: thread per entry, LDS atomic add into acc[local row]
//    (entries of one row in one block may be summed in any order).
#include "kcommon.h"

namespace rvvla {

template <class VT>
__global__ __launch_bounds__(kWG) void k_hcsr(
    const int32_t* __restrict__ brow_ptr, const int32_t* __restrict__ nb_col,
    const int32_t* __restrict__ nb_ptr, const uint32_t* __restrict__ nb_roff,
    const uint32_t* __restrict__ rp32, const uint16_t* __restrict__ coo16,
    const uint16_t* __restrict__ col16, const VT* __restrict__ val,
    const uint8_t* __restrict__ flat, const uint16_t* __restrict__ ebc16,
    const typename Val<VT>::io* __restrict__ x, typename Val<VT>::io* __restrict__ y, int64_t m,
    int R, int C, typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  extern __shared__ unsigned char lds_raw[];
  A* acc = reinterpret_cast<A*>(lds_raw);
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * R;
  const int rows = (int)((m - r0) < R ? (m - r0) : R);
  for (int i = tid; i < rows; i += kWG) acc[i] = 0;
  __syncthreads();
  const int k0 = brow_ptr[blockIdx.x], k1 = brow_ptr[blockIdx.x + 1];
  if (flat && flat[blockIdx.x]) {
    // many blocks: one pass over the block-row's entries, thread per entry
    const int E0 = nb_ptr[k0], E1 = nb_ptr[k1];
    for (int e = E0 + tid; e < E1; e += kWG)
      atomicAdd(&acc[coo16[e]],
                Val<VT>::get(val, e) * (A)x[(int64_t)ebc16[e] * C + col16[e]]);
    __syncthreads();
    for (int i = tid; i < rows; i += kWG) y[r0 + i] = alpha * acc[i] + scaled_y(beta, y, r0 + i);
    return;
  }
  for (int k = k0; k < k1; ++k) {
    const auto* xb = x + (int64_t)nb_col[k] * C;
    const int e0 = nb_ptr[k], e1 = nb_ptr[k + 1];
    const uint32_t ro = nb_roff[k];
    if ((ro >> 30) == 1) {
      const uint32_t* rp = rp32 + (ro & 0x3FFFFFFFu);
      for (int r = tid; r < rows; r += kWG) {
        const int lo = e0 + (int)rp[r], hi = e0 + (int)rp[r + 1];
        A s = 0;
        for (int e = lo; e < hi; ++e) s += Val<VT>::get(val, e) * (A)xb[col16[e]];
        acc[r] += s;
      }
    } else {
      for (int e = e0 + tid; e < e1; e += kWG)
        atomicAdd(&acc[coo16[e]], Val<VT>::get(val, e) * (A)xb[col16[e]]);
    }
    __syncthreads();
  }
  for (int i = tid; i < rows; i += kWG) y[r0 + i] = alpha * acc[i] + scaled_y(beta, y, r0 + i);
}

template <class VT>
static int run_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevHCSR& d = A.hcsr;
  if (d.b_m == 0) return RVVLA_OK;
  const size_t lds = (size_t)d.R * sizeof(AC);
  hipLaunchKernelGGL((k_hcsr<VT>), dim3((unsigned)d.b_m), dim3(kWG), lds, s, d.brow_ptr, d.nb_col,
                     d.nb_ptr, d.nb_roff, d.rp32, d.coo16, d.col16, (const VT*)d.val, d.flat,
                     d.ebc16, (const IO*)x,
                     (IO*)y, A.m, (int)d.R, (int)d.C, (AC)alpha, (AC)beta);
  return hip_check(hipGetLastError(), "k_hcsr launch");
}

int launch_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s) {
  switch (A.dtype) {
    case RVVLA_F64: return run_hcsr<double>(A, alpha, x, beta, y, s);
    case RVVLA_F32: return run_hcsr<float>(A, alpha, x, beta, y, s);
    case RVVLA_BF16: return run_hcsr<bf16_t>(A, alpha, x, beta, y, s);
    case RVVLA_F16: return run_hcsr<_Float16>(A, alpha, x, beta, y, s);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
