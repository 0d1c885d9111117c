// k_hcsr.hip -- HCSR SpMV for gfx950.
//
// Semantics follow reference src/scalar/spmv_mtx_hcsr.cpp:52-76: every
// block of block-row br adds alpha * (block . x[bc*C ...]) to y[br*R ...],
// after y has been scaled by beta; y = alpha * acc + beta * y, with y read
// and written once and only the non-empty blocks visited (the reference
// walks all b_n of them).
//  * k_hcsr_units: a block-row is cut into units of kWG local rows, one
//    workgroup each (one row per thread, so a block-row of the reference's
//    R = 2000 spreads over eight workgroups instead of one).  A CSR-kind
//    block adds the thread's row (local row_ptr from rp32, 16-bit columns
//    relative to the block's first column) into a register; a COO-kind
//    block, whose entries are ordered by local row at upload, hands the
//    unit's entry range (csplit) to the threads, one entry each, with an
//    LDS atomic add into acc[local row] (entries of one row in one block
//    may be summed in any order).  No barrier between blocks: only the
//    atomics touch acc until the final one.
//  * k_hcsr_flat: a block-row with more than kHcsrFlat non-empty blocks is
//    walked entry by entry by one workgroup (each entry's local row and
//    block column precomputed at upload), R partial sums in LDS.
#include "kcommon.h"

namespace rvvla {

template <class VT>
__global__ __launch_bounds__(kWG) void k_hcsr_units(
    const int32_t* __restrict__ u_br, const int32_t* __restrict__ u_lr0,
    const int32_t* __restrict__ brow_ptr, const int32_t* __restrict__ nb_col,
    const int32_t* __restrict__ nb_ptr, const uint32_t* __restrict__ nb_roff,
    const uint32_t* __restrict__ rp32, const int32_t* __restrict__ csplit,
    const uint16_t* __restrict__ coo16, const uint16_t* __restrict__ col16,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, int64_t m, int R, int C, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  __shared__ A acc[kWG];
  const int tid = threadIdx.x;
  const int br = u_br[blockIdx.x], lr0 = u_lr0[blockIdx.x];
  const int64_t r0 = (int64_t)br * R;
  const int rows = (int)((m - r0) < R ? (m - r0) : R);
  const int nrows = min(kWG, rows - lr0);
  const int unit = lr0 / kWG;
  acc[tid] = 0;
  __syncthreads();
  const int lr = lr0 + tid;
  const bool own = tid < nrows;
  A s = 0;
  const int k0 = brow_ptr[br], k1 = brow_ptr[br + 1];
  for (int k = k0; k < k1; ++k) {  // block descriptors: uniform (scalar) loads
    const auto* xb = x + (int64_t)nb_col[k] * C;
    const int e0 = nb_ptr[k];
    const uint32_t ro = nb_roff[k], off = ro & 0x3FFFFFFFu;
    if ((ro >> 30) == 1) {
      if (own) {
        const uint32_t* rp = rp32 + off;
        const int lo = e0 + (int)rp[lr], hi = e0 + (int)rp[lr + 1];
        for (int e = lo; e < hi; ++e) s += Val<VT>::get(val, e) * (A)xb[col16[e]];
      }
    } else {
      const int a0 = e0 + csplit[off + unit], a1 = e0 + csplit[off + unit + 1];
      for (int e = a0 + tid; e < a1; e += kWG)
        atomicAdd(&acc[coo16[e] - lr0], Val<VT>::get(val, e) * (A)xb[col16[e]]);
    }
  }
  __syncthreads();
  if (own) y[r0 + lr] = alpha * (s + acc[tid]) + scaled_y(beta, y, r0 + lr);
}

template <class VT>
__global__ __launch_bounds__(kWG) void k_hcsr_flat(
    const int32_t* __restrict__ flat_rows, const int32_t* __restrict__ brow_ptr,
    const int32_t* __restrict__ nb_ptr, const uint16_t* __restrict__ coo16,
    const uint16_t* __restrict__ col16, const VT* __restrict__ val,
    const uint16_t* __restrict__ ebc16, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, int64_t m, int R, int C, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  extern __shared__ unsigned char lds_raw[];
  A* acc = reinterpret_cast<A*>(lds_raw);
  const int tid = threadIdx.x;
  const int br = flat_rows[blockIdx.x];
  const int64_t r0 = (int64_t)br * R;
  const int rows = (int)((m - r0) < R ? (m - r0) : R);
  for (int i = tid; i < rows; i += kWG) acc[i] = 0;
  __syncthreads();
  const int E0 = nb_ptr[brow_ptr[br]], E1 = nb_ptr[brow_ptr[br + 1]];
  for (int e = E0 + tid; e < E1; e += kWG)
    atomicAdd(&acc[coo16[e]], Val<VT>::get(val, e) * (A)x[(int64_t)ebc16[e] * C + col16[e]]);
  __syncthreads();
  for (int i = tid; i < rows; i += kWG) y[r0 + i] = alpha * acc[i] + scaled_y(beta, y, r0 + i);
}

template <class VT>
static int run_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevHCSR& d = A.hcsr;
  if (d.nunits > 0) {
    hipLaunchKernelGGL((k_hcsr_units<VT>), dim3((unsigned)d.nunits), dim3(kWG), 0, s, d.u_br,
                       d.u_lr0, d.brow_ptr, d.nb_col, d.nb_ptr, d.nb_roff, d.rp32, d.csplit,
                       d.coo16, d.col16, (const VT*)d.val, (const IO*)x, (IO*)y, A.m, (int)d.R,
                       (int)d.C, (AC)alpha, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_hcsr_units launch")) return rc;
  }
  if (d.nflat > 0) {
    const size_t lds = (size_t)d.R * sizeof(AC);
    hipLaunchKernelGGL((k_hcsr_flat<VT>), dim3((unsigned)d.nflat), dim3(kWG), lds, s,
                       d.flat_rows, d.brow_ptr, d.nb_ptr, d.coo16, d.col16, (const VT*)d.val,
                       d.ebc16, (const IO*)x, (IO*)y, A.m, (int)d.R, (int)d.C, (AC)alpha,
                       (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_hcsr_flat launch")) return rc;
  }
  return RVVLA_OK;
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
