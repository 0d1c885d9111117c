// k_batch.hip -- many independent matrices, each with its own number of
// right-hand sides (1..16), in one persistent launch fed by a device work
// queue (no reference counterpart: the reference multiplies one matrix with
// one vector per call).
//
// Every matrix is cut into work items of about equal cost (a run of rows
// for the CSR item kernel, a run of 64-row slices for the SELL item
// kernel).  The matrices are packed into one queue per XCD (longest
// processing time first), so that the matrices in flight on an XCD, and
// their X, are few and stay in that XCD's L2.  A persistent grid of
// wavefronts pops items with one atomic per item: workgroup b serves queue
// b mod nqueues first, then helps with the others until all are empty.  A
// small kernel clears the counters first (a fill kernel, not a memset, so
// the pair can be captured in a graph and replayed).
//  * CSR item: a wavefront works on 64/gw rows at once, gw lanes per row.
//    The lanes of a row are (slot, rhs group) pairs: lpe lanes cover the k
//    right-hand sides of one entry, vw consecutive ones each (one 16-byte
//    or narrower load of X per entry), and the gw/lpe slots stride over the
//    row's entries; the slots are combined by a butterfly.
//  * SELL item: lane per row of a 64-row slice, k accumulators per lane.
// Y = alpha * A X + beta * Y, X n x k and Y m x k row-major per matrix with
// row stride ld (k rounded up to whole vector loads).  The padding columns
// of X are read (they hold zeros unless the caller wrote there) but only
// feed lanes whose results are dropped; those of Y are never written.
#include "kcommon.h"

namespace rvvla {

template <class T>
__global__ void k_batch_reset(int* counter, int nqueues) {
  if (blockIdx.x == 0 && (int)threadIdx.x < nqueues) counter[threadIdx.x * kQueueStride] = 0;
}

template <class T, int V>
struct VecOf {
  typedef T type __attribute__((ext_vector_type(V)));
};

template <class T, int V>
__device__ __forceinline__ void load_v(const T* __restrict__ p, T (&o)[V]) {
  if constexpr (V == 1) {
    o[0] = *p;
  } else {
    const typename VecOf<T, V>::type v = *(const typename VecOf<T, V>::type*)p;
#pragma unroll
    for (int q = 0; q < V; ++q) o[q] = v[q];
  }
}

template <class T, int V>
__device__ void batch_csr_rows(const BatchMat& M, int first, int count,
                               const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                               const T* __restrict__ val, const T* __restrict__ X,
                               T* __restrict__ Y, T alpha, T beta) {
  const int lane = threadIdx.x & 63;
  const int k = M.k, ld = M.ld, lpe = M.lpe, G = M.gw, S = G / lpe;
  const int g = lane / G, t = lane & (G - 1), slot = t / lpe, c = (t & (lpe - 1)) * V;
  const bool act = c < k;
  const int32_t* rp = rst + M.rst_off;
  const int32_t* cp = col + M.ent_off;
  const T* vp = val + M.ent_off;
  const T* xp = X + M.x_off + c;
  T* yp = Y + M.y_off + c;
  const int last = min(first + count, M.m);
  // row bounds of the next step are fetched while this step runs
  int nb0 = 0, nb1 = 0;
  if (first + g < last && act) {
    nb0 = rp[first + g];
    nb1 = rp[first + g + 1];
  }
  for (int row0 = first; row0 < last; row0 += 64 / G) {
    const int row = row0 + g;
    const int b0 = nb0, b1 = nb1;
    const int nrow = row + 64 / G;
    if (nrow < last && act) {
      nb0 = rp[nrow];
      nb1 = rp[nrow + 1];
    }
    T a[V], b[V];
#pragma unroll
    for (int q = 0; q < V; ++q) a[q] = b[q] = T(0);
    if (row < last && act) {
      const int e1 = b1;
      int e = b0 + slot;
      // four entries in flight per lane: columns and values, then X
      for (; e + 3 * S < e1; e += 4 * S) {
        const int c0 = cp[e], c1 = cp[e + S], c2 = cp[e + 2 * S], c3 = cp[e + 3 * S];
        const T v0 = vp[e], v1 = vp[e + S], v2 = vp[e + 2 * S], v3 = vp[e + 3 * S];
        T x0[V], x1[V], x2[V], x3[V];
        load_v<T, V>(xp + (int64_t)c0 * ld, x0);
        load_v<T, V>(xp + (int64_t)c1 * ld, x1);
        load_v<T, V>(xp + (int64_t)c2 * ld, x2);
        load_v<T, V>(xp + (int64_t)c3 * ld, x3);
#pragma unroll
        for (int q = 0; q < V; ++q) {
          a[q] += v0 * x0[q] + v2 * x2[q];
          b[q] += v1 * x1[q] + v3 * x3[q];
        }
      }
      for (; e < e1; e += S) {
        const T v0 = vp[e];
        T x0[V];
        load_v<T, V>(xp + (int64_t)cp[e] * ld, x0);
#pragma unroll
        for (int q = 0; q < V; ++q) a[q] += v0 * x0[q];
      }
    }
#pragma unroll
    for (int q = 0; q < V; ++q) a[q] += b[q];
    for (int off = G >> 1; off >= lpe; off >>= 1)
#pragma unroll
      for (int q = 0; q < V; ++q) a[q] += __shfl_xor(a[q], off, 64);
    if (row < last && act && slot == 0) {
      T* yo = yp + (int64_t)row * ld;
#pragma unroll
      for (int q = 0; q < V; ++q)
        if (c + q < k) yo[q] = alpha * a[q] + scaled_y(beta, yo, q);  // not the padding
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
  const int k = M.k, ld = M.ld;
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
      const T* xr = xp + (int64_t)col[idx] * ld;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < k) acc[q] += v * xr[q];
    }
    const int row = perm[M.row_off + s * 64 + lane];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (q < k) {
        T* yo = yp + (int64_t)row * ld + q;
        *yo = alpha * acc[q] + scaled_y(beta, yo, 0);
      }
  }
}

template <class T>
__device__ __forceinline__ void run_item(const BatchArgs& a, const BatchItem& it,
                                         const T* __restrict__ val, const T* __restrict__ sval,
                                         const T* __restrict__ X, T* __restrict__ Y, T alpha,
                                         T beta) {
  const BatchMat M = a.mats[it.mat];
  if (M.fmt == 0) {
    if (M.vw == 4)
      batch_csr_rows<T, 4>(M, it.first, it.count, a.rst, a.col, val, X, Y, alpha, beta);
    else if (M.vw == 2)
      batch_csr_rows<T, 2>(M, it.first, it.count, a.rst, a.col, val, X, Y, alpha, beta);
    else
      batch_csr_rows<T, 1>(M, it.first, it.count, a.rst, a.col, val, X, Y, alpha, beta);
  } else {
    batch_sell_slices<T>(M, it.first, it.count, a.cl, a.cs, a.scol, sval, a.perm, X, Y, alpha,
                         beta);
  }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_batch(BatchArgs a, const T* __restrict__ val,
                                                const T* __restrict__ sval,
                                                const T* __restrict__ X, T* __restrict__ Y,
                                                T alpha, T beta) {
  const int lane = threadIdx.x & 63;
  const int home = (int)(blockIdx.x % (unsigned)a.nqueues);
  // every wavefront leaves once it has found every queue empty
  for (int qi = 0; qi < a.nqueues; ++qi) {
    const int q = (home + qi) % a.nqueues;
    const int64_t q0 = a.qbase[q], qn = a.qbase[q + 1] - q0;
    int* cnt = a.counter + q * kQueueStride;
    for (;;) {
      int item = 0;
      if (lane == 0) item = atomicAdd(cnt, 1);
      item = __shfl(item, 0, 64);
      if (item >= qn) break;
      run_item<T>(a, a.items[q0 + item], val, sval, X, Y, alpha, beta);
    }
  }
}

template <class T>
static int run_batch(const BatchArgs& a, const void* val, const void* sval, const void* X, void* Y,
                     double alpha, double beta, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_reset<T>), dim3(1), dim3(64), 0, s, a.counter, a.nqueues);
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
