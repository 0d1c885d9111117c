// kcommon.h -- device helpers shared by the gfx950 SpMV kernels.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <cstdint>

#include "devmat.h"

namespace rvvla {

// 16-bit value storage: bf16 is kept as raw bits, fp16 as _Float16.
struct bf16_t {
  uint16_t bits;
};

template <class VT>
struct Val;
template <>
struct Val<double> {
  using acc = double;
  using io = double;
  static __device__ __forceinline__ double get(const double* p, int64_t i) { return p[i]; }
  static __device__ __forceinline__ double get_nt(const double* p, int64_t i) {
    return __builtin_nontemporal_load(p + i);
  }
};
template <>
struct Val<float> {
  using acc = float;
  using io = float;
  static __device__ __forceinline__ float get(const float* p, int64_t i) { return p[i]; }
  static __device__ __forceinline__ float get_nt(const float* p, int64_t i) {
    return __builtin_nontemporal_load(p + i);
  }
};
template <>
struct Val<bf16_t> {
  using acc = float;
  using io = float;
  static __device__ __forceinline__ float get(const bf16_t* p, int64_t i) {
    return __uint_as_float((uint32_t)p[i].bits << 16);
  }
  static __device__ __forceinline__ float get_nt(const bf16_t* p, int64_t i) {
    return __uint_as_float((uint32_t)__builtin_nontemporal_load((const uint16_t*)p + i) << 16);
  }
};
template <>
struct Val<_Float16> {
  using acc = float;
  using io = float;
  static __device__ __forceinline__ float get(const _Float16* p, int64_t i) { return (float)p[i]; }
  static __device__ __forceinline__ float get_nt(const _Float16* p, int64_t i) {
    return (float)__builtin_nontemporal_load(p + i);
  }
};

// 16-bit value storage on the device, rounded to nearest even (the host
// conversions f32_to_bf16 / f32_to_f16 give the same bits).
__device__ __forceinline__ uint16_t dev_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t dev_f16(float f) {
  const _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}

// Streamed loads: the nontemporal form marks the line as not worth keeping
// in the caches (used where x must stay resident in L2).
template <bool NT, class VT>
__device__ __forceinline__ typename Val<VT>::acc ldv(const VT* p, int64_t i) {
  if constexpr (NT) return Val<VT>::get_nt(p, i);
  else return Val<VT>::get(p, i);
}
template <bool NT>
__device__ __forceinline__ int32_t ldi(const int32_t* p, int64_t i) {
  if constexpr (NT) return __builtin_nontemporal_load(p + i);
  else return p[i];
}
template <bool NT>
__device__ __forceinline__ uint32_t ldu(const uint32_t* p, int64_t i) {
  if constexpr (NT) return __builtin_nontemporal_load(p + i);
  else return p[i];
}

// Nontemporal streaming of the matrix arrays is on unless the environment
// sets RVVLA_NT=0 (for comparisons).
static inline bool nt_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("RVVLA_NT");
    return !(e && e[0] == '0');
  }();
  return on;
}

// beta * y[i], with y not read at all when beta == 0 (BLAS convention: y
// is then write-only, so NaN/Inf already in y do not propagate).
template <class A, class IO>
__device__ __forceinline__ A scaled_y(A beta, const IO* y, int64_t i) {
  return beta != A(0) ? beta * (A)y[i] : A(0);
}

// Exact sum of two floats (Knuth's TwoSum): s + e == a + b.
template <class T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
#pragma clang fp contract(off)
  const T t = a + b;
  const T z = t - a;
  e = (a - (t - z)) + (b - z);
  s = t;
}

// Accumulator: plain, or compensated (value + running error) for long rows
// in reduced precision.
template <class T, bool COMP>
struct Acc {
  T s = 0, c = 0;
  __device__ __forceinline__ void add(T p) {
    if constexpr (COMP) {
      T e;
      two_sum(s, p, s, e);
      c += e;
    } else {
      s += p;
    }
  }
  __device__ __forceinline__ void merge(const Acc& o) {
    if constexpr (COMP) {
      T e;
      two_sum(s, o.s, s, e);
      c += o.c + e;
    } else {
      s += o.s;
    }
  }
  __device__ __forceinline__ T value() const { return COMP ? s + c : s; }
};

template <class T>
__device__ __forceinline__ T shfl_xor(T v, int m) {
  return __shfl_xor(v, m, 64);
}

// Sum of part[q * 64 + lane] over the pieces q in [q0, q1), computed by a
// whole workgroup: wavefront w adds pieces q0 + w, q0 + w + 4, ... with
// eight loads in flight, and the four wavefront sums are added in
// wavefront order through LDS, so the order of the additions is fixed by
// the piece count alone.  Every thread of the workgroup must call it; each
// gets the sum for its lane.  lds holds kWG values.
template <class A>
__device__ __forceinline__ A wg_piece_sum(const A* __restrict__ part, int64_t q0, int64_t q1,
                                          A* lds) {
  constexpr int W = kWG / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  A a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = A(0);
  int64_t q = q0 + wv;
  for (; q + 7 * W < q1; q += 8 * W) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] += part[(q + i * W) * 64 + lane];
  }
  for (; q < q1; q += W) a[0] += part[q * 64 + lane];
  lds[wv * 64 + lane] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  A s = lds[lane];
#pragma unroll
  for (int w = 1; w < W; ++w) s += lds[w * 64 + lane];
  return s;
}

// Sum over aligned groups of `width` lanes (power of two <= 64); every lane
// of the wavefront must call it.
template <class T, bool COMP>
__device__ __forceinline__ Acc<T, COMP> group_reduce(Acc<T, COMP> a, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) {
    Acc<T, COMP> o;
    o.s = shfl_xor(a.s, off);
    if constexpr (COMP) o.c = shfl_xor(a.c, off);
    a.merge(o);
  }
  return a;
}

}  // namespace rvvla
