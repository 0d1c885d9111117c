// k_batch.hip -- many independent matrices, each with its own number of
// right-hand sides (1..16), in one persistent launch fed by a device work
// queue (no reference counterpart: the reference multiplies one matrix with
// one vector per call).
//
// Every matrix is cut into work items (a run of rows for the CSR item
// kernel, a run of 64-row slices for the SELL item kernel); a persistent
// grid of wavefronts pops items from a global counter with one atomic per
// item and runs them until the queue is empty.  A small kernel clears the
// counter first (a fill kernel, not a memset, so the pair can be captured in
// a graph and replayed).
//  * CSR item: one row at a time per wavefront; lanes are (entry slot, rhs)
//    pairs with kp = next power of two >= k lanes per entry, so 64/kp entries
//    are in flight per step and each entry reads kp consecutive values of X;
//    lanes of one rhs are combined by a butterfly over the slot bits.
//  * SELL item: lane per row of a 64-row slice, k accumulators per lane.
// Y = alpha * A X + beta * Y, X n x k and Y m x k row-major per matrix.
#include "kcommon.h"

namespace rvvla {

template <class T>
__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

template <class T>
__device__ void batch_csr_rows(const BatchMat& M, int first, int count, const int32_t* __restrict__ rst,
                               const int32_t* __restrict__ col, const T* __restrict__ val,
                               const T* __restrict__ X, T* __restrict__ Y, T alpha, T beta) {
  const int lane = threadIdx.x & 63;
  const int k = M.k, kp = M.kp;
  const int slot = lane / kp, r = lane & (kp - 1), per = 64 / kp;
  const int32_t* rp = rst + M.rst_off;
  const int32_t* cp = col + M.ent_off;
  const T* vp = val + M.ent_off;
  const T* xp = X + M.x_off;
  T* yp = Y + M.y_off;
  const int last = min(first + count, M.m);
  for (int row = first; row < last; ++row) {
    const int e0 = rp[row], e1 = rp[row + 1];
    T acc = 0;
    if (r < k)
      for (int e = e0 + slot; e < e1; e += per) acc += vp[e] * xp[(int64_t)cp[e] * k + r];
    for (int off = 32; off >= kp; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (slot == 0 && r < k) {
      T* yo = yp + (int64_t)row * k + r;
      *yo = alpha * acc + scaled_y(beta, yo, 0);
    }
  }
}

template <class T>
__device__ void batch_sell_slices(const BatchMat& M, int first, int count,
                                  const int32_t* __restrict__ cl, const int32_t* __restrict__ cs,
                                  const int32_t* __restrict__ col, const T* __restrict__ val,
                                  const int32_t* __restrict__ perm, const T* __restrict__ X,
                                  T* __restrict__ Y, T alpha, T beta) {
  const int lane = threadIdx.x & 63;
  const int k = M.k;
  const T* xp = X + M.x_off;
  T* yp = Y + M.y_off;
  const int nsl = (M.m + 63) / 64;
  const int last = min(first + count, nsl);
  for (int s = first; s < last; ++s) {
    const int height = min(64, M.m - s * 64);
    if (lane >= height) continue;
    const int width = cl[M.sl_off + s];
    int64_t idx = M.slot_off + cs[M.sl_off + s] + lane;
    T acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0;
    for (int j = 0; j < width; ++j, idx += height) {
      const T v = val[idx];
      const T* xr = xp + (int64_t)col[idx] * k;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < k) acc[q] += v * xr[q];
    }
    const int row = perm[M.row_off + s * 64 + lane];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (q < k) {
        T* yo = yp + (int64_t)row * k + q;
        *yo = alpha * acc[q] + scaled_y(beta, yo, 0);
      }
  }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_batch(BatchArgs a, const T* __restrict__ val,
                                                const T* __restrict__ sval,
                                                const T* __restrict__ X, T* __restrict__ Y,
                                                T alpha, T beta) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(a.counter, 1);
    item = __shfl(item, 0, 64);
    if (item >= a.nitems) return;  // every wavefront leaves once the queue is empty
    const BatchItem it = a.items[item];
    const BatchMat M = a.mats[it.mat];
    if (M.fmt == 0)
      batch_csr_rows<T>(M, it.first, it.count, a.rst, a.col, val, X, Y, alpha, beta);
    else
      batch_sell_slices<T>(M, it.first, it.count, a.cl, a.cs, a.scol, sval, a.perm, X, Y, alpha,
                           beta);
  }
}

template <class T>
static int run_batch(const BatchArgs& a, const void* val, const void* sval, const void* X, void* Y,
                     double alpha, double beta, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_reset<T>), dim3(1), dim3(64), 0, s, a.counter);
  if (int rc = hip_check(hipGetLastError(), "k_batch_reset launch")) return rc;
  if (a.nitems == 0) return RVVLA_OK;
  hipLaunchKernelGGL((k_batch<T>), dim3((unsigned)grid), dim3(kWG), 0, s, a, (const T*)val,
                     (const T*)sval, (const T*)X, (T*)Y, (T)alpha, (T)beta);
  return hip_check(hipGetLastError(), "k_batch launch");
}

int launch_batch(int dtype, const BatchArgs& a, const void* val, const void* sval, const void* X,
                 void* Y, double alpha, double beta, int grid, hipStream_t s) {
  if (dtype == RVVLA_F64) return run_batch<double>(a, val, sval, X, Y, alpha, beta, grid, s);
  if (dtype == RVVLA_F32) return run_batch<float>(a, val, sval, X, Y, alpha, beta, grid, s);
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
