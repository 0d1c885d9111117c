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

// CUs of the current device (256 when it cannot be read), looked up once
// per device.
static inline int device_cu_count() {
  static int cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = __atomic_load_n(&cache[dev], __ATOMIC_RELAXED);
  if (c == 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
      c = 256;
    __atomic_store_n(&cache[dev], c, __ATOMIC_RELAXED);
  }
  return c;
}

// XCD-contiguous work order.  The dispatcher deals workgroup b to XCD
// b mod 8; for a grid of G workgroups, G a multiple of 8 (launchers round
// up, and the extra workgroups find no work), this renumbering gives XCD k
// the contiguous run [k G/8, (k+1) G/8), so that neighbouring work units,
// which share lines of x and of the streamed arrays, meet in one L2.
// Measured on the 2048^2 Laplacian it pays for 32-bit accumulation (VNEC,
// VHCC, CVR, CSR5, HCSR: 5-15 %) and is neutral or costs FP64 up to 10 %,
// so the launchers use it for 32-bit accumulators only.
__device__ __forceinline__ int xcd_run_index(int b, int G) { return (b & 7) * (G >> 3) + (b >> 3); }
constexpr int64_t xcd_grid(int64_t g) { return (g + 7) / 8 * 8; }

// y[i] = v, FP64 results with a nontemporal store.  Used where a
// wavefront writes whole runs of consecutive rows (SELL, CVR, fused LAV):
// measured 1-3 % faster there on the 2048^2 Laplacian, and slower where
// lanes write scattered rows (CSR5 +25 %, VHCC +4 %) or in the LDS-sliced
// combine (+2 %).
template <class IO, class A>
__device__ __forceinline__ void put_y(IO* __restrict__ y, int64_t i, A v) {
  const IO o = (IO)v;
  if constexpr (sizeof(IO) == 8) __builtin_nontemporal_store(o, y + i);
  else y[i] = o;
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

// Marks a loaded value as used here, unconditionally, so that the
// compiler cannot move its load under a later select or branch (a load
// moved into a conditional block is waited for on its own, which
// serialises a batch of gathers).  Emits no instruction.
template <class T>
__device__ __forceinline__ void keep_live(T& v) {
  asm volatile("" : "+v"(v));
}

template <class T>
__device__ __forceinline__ T shfl_xor(T v, int m) {
  return __shfl_xor(v, m, 64);
}

// Sum of part[q * 64 + lane] over the pieces q in [q0, q1), computed by a
// whole workgroup: wavefront w adds pieces q0 + w, q0 + w + 4, ... with
// eight loads in flight, and the four wavefront sums are added in
// wavefront order through LDS, so the order of the additions is fixed by
// the piece count alone.  Every thread of the workgroup must call it; each
// gets the sum for its lane.  lds holds kWG values, or 2 kWG with COMP:
// then every addition is compensated (TwoSum), for the 16-bit value forms.
template <class A, bool COMP = false>
__device__ __forceinline__ A wg_piece_sum(const A* __restrict__ part, int64_t q0, int64_t q1,
                                          A* lds) {
  constexpr int W = kWG / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  Acc<A, COMP> a[8];
  int64_t q = q0 + wv;
  for (; q + 7 * W < q1; q += 8 * W) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i].add(part[(q + i * W) * 64 + lane]);
  }
  for (; q < q1; q += W) a[0].add(part[q * 64 + lane]);
  a[0].merge(a[1]);
  a[2].merge(a[3]);
  a[4].merge(a[5]);
  a[6].merge(a[7]);
  a[0].merge(a[2]);
  a[4].merge(a[6]);
  a[0].merge(a[4]);
  lds[wv * 64 + lane] = a[0].s;
  if constexpr (COMP) lds[kWG + wv * 64 + lane] = a[0].c;
  __syncthreads();
  Acc<A, COMP> s;
  s.s = lds[lane];
  if constexpr (COMP) s.c = lds[kWG + lane];
#pragma unroll
  for (int w = 1; w < W; ++w) {
    Acc<A, COMP> o;
    o.s = lds[w * 64 + lane];
    if constexpr (COMP) o.c = lds[kWG + w * 64 + lane];
    s.merge(o);
  }
  return s.value();
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

// DPP move of a 32- or 64-bit value (CTRL: row_shr:n = 0x110 | n, wave_shr:1
// = 0x138); lanes without a source lane get 0.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float lane_read(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <class Ac>
__device__ __forceinline__ Ac acc_readlane(const Ac& a, int l) {
  Ac r;
  r.s = lane_read(a.s, l);
  r.c = lane_read(a.c, l);
  return r;
}
// One in-row Hillis-Steele step of the segmented scan: lane l adds lane
// l - D of its row when that lane is in its segment.
template <int CTRL, int D, class Ac>
__device__ __forceinline__ void seg_scan_step(Ac& S, int lr, int lane, int st) {
  Ac o;
  o.s = dpp_mov<CTRL>(S.s);
  o.c = dpp_mov<CTRL>(S.c);
  if (lr >= D && lane - D >= st) S.merge(o);
}

// Segmented inclusive scan of a across the wavefront: lane l gets the sum
// over lanes st..l, st the highest lane <= l whose bit is set in heads
// (lane 0's bit must be set).  Hillis-Steele steps inside each row of 16
// lanes (DPP row shifts), then the open sum at the end of each row carried
// into the next rows' lanes whose segment started before them; the order
// of the additions is fixed by heads alone.
template <class Ac>
__device__ __forceinline__ Ac wave_seg_scan(Ac S, uint64_t heads, int lane) {
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int st = 63 - __clzll((long long)(heads & upto));
  const int lr = lane & 15;
  seg_scan_step<0x111, 1>(S, lr, lane, st);
  seg_scan_step<0x112, 2>(S, lr, lane, st);
  seg_scan_step<0x114, 4>(S, lr, lane, st);
  seg_scan_step<0x118, 8>(S, lr, lane, st);
  // row r's lanes take the open sum at the end of row r - 1, which already
  // holds the carries of the rows before it
  const Ac T = acc_readlane(S, 15);
  if (lane >= 16 && lane < 32 && st < 16) S.merge(T);
  const Ac E = acc_readlane(S, 31);
  if (lane >= 32 && lane < 48 && st < 32) S.merge(E);
  const Ac F = acc_readlane(S, 47);
  if (lane >= 48 && st < 48) S.merge(F);
  return S;
}

}  // namespace rvvla
