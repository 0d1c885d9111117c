// k_grad.hip -- the kernels behind differentiable SpMM (no reference
// counterpart: the reference multiplies one vector with fixed values).
//
//  * k_spmm_vec: Y = alpha * A(vals) X + beta * Y with the CSR structure of
//    A and values from the caller (e.g. a parameter tensor), 16-byte
//    vectors: one wavefront per row, a group of lanes per entry (each lane
//    one vector of the entry's k-row of X) and several entries' gathers in
//    flight; k_spmm_vals (lane j owns column j) serves widths that do not
//    split into vectors.
//  * k_spmm_t: Y = alpha * A(vals)^T X + beta * Y over the cached CSC
//    structure: wavefront per column of A, values gathered through the CSC
//    -> CSR permutation, no atomics.
//  * k_sddmm: out[e] = alpha * <DY[row_e, :], X[col_e, :]> (the gradient of
//    the values), 16 lanes per entry with float4/double2-wide partial dots
//    and a 4-step butterfly, four entries per wavefront step and four steps'
//    gathers in flight.
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

__device__ __forceinline__ float4 vaxpy(float a, float4 x, float4 acc) {
  return make_float4(acc.x + a * x.x, acc.y + a * x.y, acc.z + a * x.z, acc.w + a * x.w);
}
__device__ __forceinline__ double2 vaxpy(double a, double2 x, double2 acc) {
  return make_double2(acc.x + a * x.x, acc.y + a * x.y);
}
__device__ __forceinline__ float4 vadd(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float4 vshfl_xor(float4 v, int m) {
  return make_float4(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64), __shfl_xor(v.z, m, 64),
                     __shfl_xor(v.w, m, 64));
}
__device__ __forceinline__ double2 vshfl_xor(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

// Y = alpha * A(vals) X + beta * Y with 16-byte vectors: one wavefront per
// row, L lanes per entry (lane l of a group holds vector l of the k-row),
// 64 / L entries per step in groups and U steps loaded before any is added,
// so 64 / L * U gathers of X rows are in flight per wavefront.  The groups'
// sums are then added in a fixed butterfly.  k / width <= L vectors; lanes
// past them idle.
template <class T, int L, int U>
__global__ __launch_bounds__(kWG) void k_spmm_vec(const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ idx,
                                                   const int32_t* __restrict__ perm,
                                                   const T* __restrict__ vals,
                                                   const T* __restrict__ X, int64_t ldx,
                                                   T* __restrict__ Y, int64_t ldy, int64_t rows,
                                                   int nv, T alpha, T beta) {
  using V = typename Vec4<T>::type;
  constexpr int W = sizeof(V) / sizeof(T);
  constexpr int G = 64 / L;
  const int64_t r = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)));
  if (r >= rows) return;
  const int lane = threadIdx.x & 63, g = lane / L, l = lane % L;
  const bool on = l < nv;
  const int lv = on ? l : 0;  // idle lanes gather vector 0 and add nothing
  const int e0 = ptr[r], e1 = ptr[r + 1];
  V acc = {};
  for (int eb = e0; eb < e1; eb += G * U) {
    int j[U];
    T v[U];
    V xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = min(eb + u * G + g, e1 - 1);  // past the row: repeat its last entry
      j[u] = idx[e];
      v[u] = vals[perm ? perm[e] : e];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = *reinterpret_cast<const V*>(&X[(int64_t)j[u] * ldx + lv * W]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      keep_live(xv[u].x);
      acc = vaxpy(eb + u * G + g < e1 ? v[u] : T(0), xv[u], acc);
    }
  }
#pragma unroll
  for (int off = L; off < 64; off <<= 1) acc = vadd(acc, vshfl_xor(acc, off));
  if (g == 0 && on) {
    V* yp = reinterpret_cast<V*>(&Y[r * ldy + l * W]);
    V out;
    if (beta != T(0)) {
      const V y0 = *yp;
      if constexpr (W == 4) out = make_float4(alpha * acc.x + beta * y0.x, alpha * acc.y + beta * y0.y,
                                              alpha * acc.z + beta * y0.z, alpha * acc.w + beta * y0.w);
      else out = make_double2(alpha * acc.x + beta * y0.x, alpha * acc.y + beta * y0.y);
    } else {
      if constexpr (W == 4) out = make_float4(alpha * acc.x, alpha * acc.y, alpha * acc.z, alpha * acc.w);
      else out = make_double2(alpha * acc.x, alpha * acc.y);
    }
    *yp = out;
  }
}

// The backward of Y = A(vals) X in one traversal of A's columns (the
// cached CSC structure): column c's wavefront loads X[c, :] once and, for
// every entry (r, c) of the column (CSR position p = perm[e]), gathers
// DY[r, :] once and uses it twice: DX[c, :] += vals[p] * DY[r, :] (DX) and
// dvals[p] = <DY[r, :], X[c, :]> (DV).  L lanes per entry as in k_spmm_vec;
// the column's sums are added in a fixed butterfly, the dots in one per
// entry.  DX = A^T DY (beta = 0) and dvals need no atomics: a column is one
// wavefront's and every entry is written once.
template <class T, int L, int U, bool DX, bool DV>
__global__ __launch_bounds__(kWG) void k_spmm_bwd(const int32_t* __restrict__ cptr,
                                                   const int32_t* __restrict__ crow,
                                                   const int32_t* __restrict__ perm,
                                                   const T* __restrict__ vals,
                                                   const T* __restrict__ DY, int64_t lddy,
                                                   const T* __restrict__ X, int64_t ldx,
                                                   T* __restrict__ dX, int64_t lddx,
                                                   T* __restrict__ dvals, int64_t ncols, int nv) {
  using V = typename Vec4<T>::type;
  constexpr int W = sizeof(V) / sizeof(T);
  constexpr int G = 64 / L;
  const int64_t c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)));
  if (c >= ncols) return;
  const int lane = threadIdx.x & 63, g = lane / L, l = lane % L;
  const bool on = l < nv;
  const int lv = on ? l : 0;  // idle lanes gather vector 0 and add nothing
  V xc = {};
  if constexpr (DV) xc = *reinterpret_cast<const V*>(&X[c * ldx + lv * W]);
  const int e0 = cptr[c], e1 = cptr[c + 1];
  V acc = {};
  for (int eb = e0; eb < e1; eb += G * U) {
    int p[U];
    T th[U];
    V dy[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = min(eb + u * G + g, e1 - 1);  // past the column: repeat its last entry
      p[u] = perm[e];
      dy[u] = *reinterpret_cast<const V*>(&DY[(int64_t)crow[e] * lddy + lv * W]);
    }
    if constexpr (DX) {
#pragma unroll
      for (int u = 0; u < U; ++u) th[u] = vals[p[u]];
    }
    T d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      keep_live(dy[u].x);
      const bool valid = eb + u * G + g < e1;
      if constexpr (DX) acc = vaxpy(valid ? th[u] : T(0), dy[u], acc);
      if constexpr (DV) d[u] = on ? Vec4<T>::dot(dy[u], xc) : T(0);
    }
    if constexpr (DV) {
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1)
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] += __shfl_xor(d[u], off, 64);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (l == 0 && eb + u * G + g < e1) dvals[p[u]] = d[u];
    }
  }
  if constexpr (DX) {
#pragma unroll
    for (int off = L; off < 64; off <<= 1) acc = vadd(acc, vshfl_xor(acc, off));
    if (g == 0 && on) *reinterpret_cast<V*>(&dX[c * lddx + l * W]) = acc;
  }
}

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
  constexpr int U = 4;  // entries per group loaded before any dot
  const bool on = l < nv;
  const int lv = on ? l : 0;  // idle lanes gather vector 0 and dot nothing
  const V dy = *reinterpret_cast<const V*>(&DY[r * lddy + lv * W]);
  const int e0 = rst[r], e1 = rst[r + 1];
  for (int eb = e0; eb < e1; eb += 4 * U) {
    V xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = min(eb + 4 * u + g, e1 - 1);  // past the row: repeat its last entry
      xv[u] = *reinterpret_cast<const V*>(&X[(int64_t)col[e] * ldx + lv * W]);
    }
    T p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      keep_live(xv[u].x);
      p[u] = on ? Vec4<T>::dot(dy, xv[u]) : T(0);
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1)
#pragma unroll
      for (int u = 0; u < U; ++u) p[u] += __shfl_xor(p[u], off, 64);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = eb + 4 * u + g;
      if (l == 0 && e < e1) out[e] = alpha * p[u];
    }
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
  using V = typename Vec4<T>::type;
  constexpr int W = sizeof(V) / sizeof(T);
  const int nv = k / W;
  const bool vec_ok = k % W == 0 && nv <= 32 && ldx % W == 0 && ldy % W == 0 &&
                      (uintptr_t)X % sizeof(V) == 0 && (uintptr_t)Y % sizeof(V) == 0;
  if (vec_ok) {
#define RVVLA_SPMM_VEC(L_, U_)                                                                   \
  hipLaunchKernelGGL((k_spmm_vec<T, L_, U_>), dim3(grid), dim3(kWG), 0, s, ptr, idx, perm,        \
                     (const T*)vals, (const T*)X, ldx, (T*)Y, ldy, rows, nv, (T)alpha, (T)beta)
    if (nv <= 4) RVVLA_SPMM_VEC(4, 2);
    else if (nv <= 8) RVVLA_SPMM_VEC(8, 4);
    else if (nv <= 16) RVVLA_SPMM_VEC(16, 4);
    else RVVLA_SPMM_VEC(32, 8);
#undef RVVLA_SPMM_VEC
    return hip_check(hipGetLastError(), "k_spmm_vec launch");
  }
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

template <class T>
static int run_spmm_bwd(const rvvla_mat_s& A, const void* vals, int k, const void* DY,
                        int64_t lddy, const void* X, int64_t ldx, void* dX, int64_t lddx,
                        void* dvals, hipStream_t s) {
  using V = typename Vec4<T>::type;
  constexpr int W = sizeof(V) / sizeof(T);
  const int nv = k / W;
  const bool vec_ok = k > 0 && k % W == 0 && nv <= 16 && lddy % W == 0 && ldx % W == 0 &&
                      lddx % W == 0 && (uintptr_t)DY % sizeof(V) == 0 &&
                      (uintptr_t)X % sizeof(V) == 0 && (uintptr_t)dX % sizeof(V) == 0;
  if (!vec_ok) {  // the two separate passes
    if (dX) {
      if (int rc = run_spmm_vals<T>(A, true, vals, k, 1.0, DY, lddy, 0.0, dX, lddx, s)) return rc;
    }
    return dvals ? run_sddmm<T>(A, k, DY, lddy, X, ldx, 1.0, dvals, s) : RVVLA_OK;
  }
  if (A.n == 0) return RVVLA_OK;
  const unsigned grid = (unsigned)((A.n + kWG / 64 - 1) / (kWG / 64));
#define RVVLA_BWD(L_, U_, DX_, DV_)                                                             \
  hipLaunchKernelGGL((k_spmm_bwd<T, L_, U_, DX_, DV_>), dim3(grid), dim3(kWG), 0, s, A.csc.ptr,  \
                     A.csc.row, A.csc.perm, (const T*)vals, (const T*)DY, lddy, (const T*)X, ldx, \
                     (T*)dX, lddx, (T*)dvals, A.n, nv)
#define RVVLA_BWD_L(L_, U_)                                \
  if (dX && dvals) RVVLA_BWD(L_, U_, true, true);          \
  else if (dX) RVVLA_BWD(L_, U_, true, false);             \
  else if (dvals) RVVLA_BWD(L_, U_, false, true)
  if (nv <= 4) {
    RVVLA_BWD_L(4, 2);
  } else if (nv <= 8) {
    RVVLA_BWD_L(8, 4);
  } else {
    RVVLA_BWD_L(16, 4);
  }
#undef RVVLA_BWD_L
#undef RVVLA_BWD
  return hip_check(hipGetLastError(), "k_spmm_bwd launch");
}

int launch_spmm_backward(const rvvla_mat_s& A, const void* vals, int k, const void* DY,
                         int64_t lddy, const void* X, int64_t ldx, void* dX, int64_t lddx,
                         void* dvals, hipStream_t s) {
  if (A.dtype == RVVLA_F64)
    return run_spmm_bwd<double>(A, vals, k, DY, lddy, X, ldx, dX, lddx, dvals, s);
  if (A.dtype == RVVLA_F32)
    return run_spmm_bwd<float>(A, vals, k, DY, lddy, X, ldx, dX, lddx, dvals, s);
  return RVVLA_ERR_VALUE_TYPE;
}

int launch_sddmm(const rvvla_mat_s& A, int k, const void* DY, int64_t lddy, const void* X,
                 int64_t ldx, double alpha, void* out, hipStream_t s) {
  if (A.dtype == RVVLA_F64) return run_sddmm<double>(A, k, DY, lddy, X, ldx, alpha, out, s);
  if (A.dtype == RVVLA_F32) return run_sddmm<float>(A, k, DY, lddy, X, ldx, alpha, out, s);
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
