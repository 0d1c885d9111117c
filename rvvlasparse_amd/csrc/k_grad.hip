// k_grad.hip -- the kernels behind differentiable SpMM (no reference
// counterpart: the reference multiplies one vector with fixed values).
//
//  * k_spmm_vals: Y = alpha * A(vals) X + beta * Y with the CSR structure of
//    A and values from the caller (e.g. a parameter tensor).  One wavefront
//    per row, lane j owns column j of X/Y, so every entry reads one
//    contiguous k-vector of X.
//  * k_spmm_t: Y = alpha * A(vals)^T X + beta * Y over the cached CSC
//    structure: wavefront per column of A, values gathered through the CSC
//    -> CSR permutation, no atomics.
//  * k_sddmm: out[e] = alpha * <DY[row_e, :], X[col_e, :]> (the gradient of
//    the values), 16 lanes per entry with float4/double2-wide partial dots
//    and a 4-step butterfly, four entries per wavefront step.
#include "kcommon.h"

namespace rvvla {

template <class T>
__global__ __launch_bounds__(kWG) void k_spmm_vals(const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx,
                                                    const int32_t* __restrict__ perm,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ X, int64_t ldx,
                                                    T* __restrict__ Y, int64_t ldy, int64_t rows,
                                                    int k, T alpha, T beta) {
  const int64_t r = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  const int e0 = ptr[r], e1 = ptr[r + 1];
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int c = c0 + lane;
    const bool on = c < k;
    T s0 = 0, s1 = 0;
    int e = e0;
    for (; e + 1 < e1; e += 2) {
      const int j0 = idx[e], j1 = idx[e + 1];
      const T v0 = vals[perm ? perm[e] : e], v1 = vals[perm ? perm[e + 1] : e + 1];
      if (on) {
        s0 += v0 * X[(int64_t)j0 * ldx + c];
        s1 += v1 * X[(int64_t)j1 * ldx + c];
      }
    }
    if (e < e1 && on) s0 += vals[perm ? perm[e] : e] * X[(int64_t)idx[e] * ldx + c];
    if (on) {
      T* yp = &Y[r * ldy + c];
      *yp = alpha * (s0 + s1) + scaled_y(beta, yp, 0);
    }
  }
}

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  static __device__ __forceinline__ float dot(float4 a, float4 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  }
};
template <>
struct Vec4<double> {
  using type = double2;
  static __device__ __forceinline__ double dot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
};

// k must be a multiple of the vector width (4 floats / 2 doubles) and at
// most 16 vectors (64 floats / 32 doubles); wider k uses k_sddmm_any.
template <class T>
__global__ __launch_bounds__(kWG) void k_sddmm(const int32_t* __restrict__ rst,
                                                const int32_t* __restrict__ col,
                                                const T* __restrict__ DY, int64_t lddy,
                                                const T* __restrict__ X, int64_t ldx,
                                                T* __restrict__ out, int64_t m, int k, T alpha) {
  using V = typename Vec4<T>::type;
  constexpr int W = sizeof(V) / sizeof(T);
  const int64_t r = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= m) return;
  const int g = lane >> 4, l = lane & 15;  // 4 groups of 16 lanes
  const int nv = k / W;                    // vectors per row
  V dy;
  const bool on = l < nv;
  if (on) dy = *reinterpret_cast<const V*>(&DY[r * lddy + l * W]);
  const int e0 = rst[r], e1 = rst[r + 1];
  for (int eb = e0; eb < e1; eb += 4) {
    const int e = eb + g;
    T p = 0;
    if (on && e < e1) {
      const V xv = *reinterpret_cast<const V*>(&X[(int64_t)col[e] * ldx + l * W]);
      p = Vec4<T>::dot(dy, xv);
    }
    for (int off = 8; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
    if (l == 0 && e < e1) out[e] = alpha * p;
  }
}

// General k: one wavefront per row, lanes over k, one butterfly per entry.
template <class T>
__global__ __launch_bounds__(kWG) void k_sddmm_any(const int32_t* __restrict__ rst,
                                                    const int32_t* __restrict__ col,
                                                    const T* __restrict__ DY, int64_t lddy,
                                                    const T* __restrict__ X, int64_t ldx,
                                                    T* __restrict__ out, int64_t m, int k,
                                                    T alpha) {
  const int64_t r = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= m) return;
  const int e0 = rst[r], e1 = rst[r + 1];
  for (int e = e0; e < e1; ++e) {
    T p = 0;
    for (int c = lane; c < k; c += 64) p += DY[r * lddy + c] * X[(int64_t)col[e] * ldx + c];
    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
    if (lane == 0) out[e] = alpha * p;
  }
}

template <class T>
static int run_spmm_vals(const rvvla_mat_s& A, bool transpose, const void* vals, int k,
                         double alpha, const void* X, int64_t ldx, double beta, void* Y,
                         int64_t ldy, hipStream_t s) {
  const int64_t rows = transpose ? A.n : A.m;
  if (rows == 0 || k == 0) return RVVLA_OK;
  const unsigned grid = (unsigned)((rows + kWG / 64 - 1) / (kWG / 64));
  const int32_t* ptr = transpose ? A.csc.ptr : A.csr.rst;
  const int32_t* idx = transpose ? A.csc.row : A.csr.col;
  const int32_t* perm = transpose ? A.csc.perm : nullptr;
  hipLaunchKernelGGL((k_spmm_vals<T>), dim3(grid), dim3(kWG), 0, s, ptr, idx, perm,
                     (const T*)vals, (const T*)X, ldx, (T*)Y, ldy, rows, k, (T)alpha, (T)beta);
  return hip_check(hipGetLastError(), "k_spmm_vals launch");
}

int launch_spmm_values(const rvvla_mat_s& A, bool transpose, const void* vals, int k, double alpha,
                       const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
                       hipStream_t s) {
  if (A.dtype == RVVLA_F64)
    return run_spmm_vals<double>(A, transpose, vals, k, alpha, X, ldx, beta, Y, ldy, s);
  if (A.dtype == RVVLA_F32)
    return run_spmm_vals<float>(A, transpose, vals, k, alpha, X, ldx, beta, Y, ldy, s);
  return RVVLA_ERR_VALUE_TYPE;
}

template <class T>
static int run_sddmm(const rvvla_mat_s& A, int k, const void* DY, int64_t lddy, const void* X,
                     int64_t ldx, double alpha, void* out, hipStream_t s) {
  if (A.m == 0 || A.nnz == 0) return RVVLA_OK;
  const unsigned grid = (unsigned)((A.m + kWG / 64 - 1) / (kWG / 64));
  constexpr int W = sizeof(typename Vec4<T>::type) / sizeof(T);
  const bool vec_ok = k % W == 0 && k / W <= 16 && lddy % W == 0 && ldx % W == 0 &&
                      (uintptr_t)DY % sizeof(typename Vec4<T>::type) == 0 &&
                      (uintptr_t)X % sizeof(typename Vec4<T>::type) == 0;
  if (vec_ok)
    hipLaunchKernelGGL((k_sddmm<T>), dim3(grid), dim3(kWG), 0, s, A.csr.rst, A.csr.col,
                       (const T*)DY, lddy, (const T*)X, ldx, (T*)out, A.m, k, (T)alpha);
  else
    hipLaunchKernelGGL((k_sddmm_any<T>), dim3(grid), dim3(kWG), 0, s, A.csr.rst, A.csr.col,
                       (const T*)DY, lddy, (const T*)X, ldx, (T*)out, A.m, k, (T)alpha);
  return hip_check(hipGetLastError(), "k_sddmm launch");
}

int launch_sddmm(const rvvla_mat_s& A, int k, const void* DY, int64_t lddy, const void* X,
                 int64_t ldx, double alpha, void* out, hipStream_t s) {
  if (A.dtype == RVVLA_F64) return run_sddmm<double>(A, k, DY, lddy, X, ldx, alpha, out, s);
  if (A.dtype == RVVLA_F32) return run_sddmm<float>(A, k, DY, lddy, X, ldx, alpha, out, s);
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
