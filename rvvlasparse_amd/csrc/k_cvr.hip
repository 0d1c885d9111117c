// k_cvr.hip -- CVR SpMV for gfx950 (SIMD_LEN = 64: one stream per
// wavefront, one CVR lane per wavefront lane).
//
// Semantics follow the reference scalar kernel (reference
// src/scalar/spmv_mtx_cvr.cpp:32-66): every lane accumulates the entries of
// its current row in pack order; when a record says the lane switches rows
// before pack k, the lane's sum is flushed into that row; at the stream's
// end the remaining lanes flush into their open rows.  A row flushed once
// is written directly, y = alpha*acc + beta*y (the reference scales y by
// beta first); a long row cut into segments is flushed once per segment,
// each flush into its own slot of a partial-sum buffer, and k_cvr_combine
// adds a row's slots in order; empty rows come from a separate list.  The
// records are re-laid at upload as one 64-bit flush mask per pack (a
// uniform scalar load) plus the flush targets in record order; a lane finds
// its target by counting the flushing lanes below it.  With 32-bit
// accumulation each XCD walks a contiguous run of streams.
#include "kcommon.h"

namespace rvvla {

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_cvr(
    const int64_t* __restrict__ pack_start, const int32_t* __restrict__ rec_start,
    const uint64_t* __restrict__ flush, const int32_t* __restrict__ rec_rows,
    const int32_t* __restrict__ tail_rows, const uint32_t* __restrict__ col,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc* __restrict__ part,
    int64_t nstreams, typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t t = __builtin_amdgcn_readfirstlane((int)((sizeof(A) < 8 ? xcd_run_index((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x) * (kWG / 64) + (threadIdx.x >> 6)));
  if (t >= nstreams) return;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  const int64_t k0 = pack_start[t], k1 = pack_start[t + 1];
  int32_t rb = rec_start[t];
  Acc<A, COMP> acc;  // 16-bit values (COMP): TwoSum compensated
  auto write = [&](int32_t r) {
    if (r >= 0) put_y(y, r, alpha * acc.value() + scaled_y(beta, y, r));
    else part[-(int64_t)r - 2] = acc.value();
  };
  // kBatch packs at a time: masks, flushed rows, columns, values and the
  // gathers are all issued before the packs are walked
  constexpr int kBatch = 8;
  for (int64_t kb = k0; kb < k1; kb += kBatch) {
    uint64_t mk[kBatch];
    uint32_t c[kBatch];
    A v[kBatch], p[kBatch];
    int32_t rr[kBatch];
    if (kb + kBatch <= k1) {  // a whole batch: straight-line loads
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int64_t e = (kb + u) * 64 + lane;
        mk[u] = flush[kb + u];
        c[u] = col[e];
        v[u] = Val<VT>::get(val, e);
      }
    } else {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const bool in = kb + u < k1;
        const int64_t e = (kb + u) * 64 + lane;
        mk[u] = in ? flush[kb + u] : 0ull;
        c[u] = in ? col[e] : 0u;
        v[u] = in ? Val<VT>::get(val, e) : A(0);
      }
    }
    // a pack past the stream's end adds zero (its x gather is column 0,
    // which may hold Inf): the gathers are all taken first (the empty asm),
    // so that the select does not move any of them under a branch
    A g[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      g[u] = (A)x[c[u]];
      rr[u] = ((mk[u] >> lane) & 1u) ? rec_rows[rb + __popcll(mk[u] & below)] : 0;
      rb += __popcll(mk[u]);
    }
    static_assert(kBatch == 8, "eight gathers kept live");
    __asm__ volatile("" : "+v"(g[0]), "+v"(g[1]), "+v"(g[2]), "+v"(g[3]), "+v"(g[4]), "+v"(g[5]),
                     "+v"(g[6]), "+v"(g[7]));
#pragma unroll
    for (int u = 0; u < kBatch; ++u) p[u] = kb + u < k1 ? v[u] * g[u] : A(0);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if ((mk[u] >> lane) & 1u) {
        write(rr[u]);
        acc = Acc<A, COMP>();
      }
      acc.add(p[u]);
    }
  }
  const int32_t r = tail_rows[t * 64 + lane];
  if (r != -1) write(r);
}

// Rows flushed more than once: y = alpha * (their slots' sum, in slot
// order) + beta * y.  Rows with more than 64 slots take a wavefront each
// (lane l adds slots l, l + 64, ...; the lane sums are then added in a
// fixed tree), the others a thread each.
template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_cvr_combine_long(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ ptr, int64_t n,
    const typename Val<VT>::acc* __restrict__ part, typename Val<VT>::io* __restrict__ y,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t i = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)));
  if (i >= n) return;
  const int lane = threadIdx.x & 63;
  Acc<A, COMP> s;
  for (int64_t q = ptr[2 * i] + lane; q < ptr[2 * i + 1]; q += 64) s.add(part[q]);
  s = group_reduce(s, 64);
  if (lane == 0) {
    const int32_t r = rows[i];
    y[r] = alpha * s.value() + scaled_y(beta, y, r);
  }
}

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_cvr_combine(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ ptr, int64_t i0, int64_t n,
    const typename Val<VT>::acc* __restrict__ part, typename Val<VT>::io* __restrict__ y,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t i = i0 + (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i >= n) return;
  Acc<A, COMP> s;
  for (int64_t q = ptr[2 * i]; q < ptr[2 * i + 1]; ++q) s.add(part[q]);
  const int32_t r = rows[i];
  y[r] = alpha * s.value() + scaled_y(beta, y, r);
}

template <class VT>
__global__ __launch_bounds__(kWG) void k_scale_rows(const int32_t* __restrict__ rows, int64_t n,
                                                     typename Val<VT>::io* __restrict__ y,
                                                     typename Val<VT>::acc beta) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < n) {
    const int32_t r = rows[i];
    y[r] = scaled_y(beta, y, r);
  }
}

template <class VT>
static int run_cvr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                   hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  constexpr bool COMP = sizeof(VT) == 2;
  const DevCVR& d = A.cvr;
  AC* part = (AC*)(scratch ? scratch : d.part);
  if (d.nstreams > 0) {
    const unsigned g0 = (unsigned)((d.nstreams + kWG / 64 - 1) / (kWG / 64));
    const unsigned g = sizeof(AC) < 8 ? (unsigned)xcd_grid(g0) : g0;
    hipLaunchKernelGGL((k_cvr<VT, COMP>), dim3(g), dim3(kWG), 0, s, d.pack_start, d.rec_start,
                       d.flush, d.rec_rows, d.tail_rows, d.col, (const VT*)d.val, (const IO*)x,
                       (IO*)y, part, d.nstreams, (AC)alpha, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_cvr launch")) return rc;
  }
  if (d.nsplit_long > 0) {
    const unsigned g = (unsigned)((d.nsplit_long + kWG / 64 - 1) / (kWG / 64));
    hipLaunchKernelGGL((k_cvr_combine_long<VT, COMP>), dim3(g), dim3(kWG), 0, s, d.split_rows,
                       d.split_ptr, d.nsplit_long, (const AC*)part, (IO*)y, (AC)alpha, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_cvr_combine_long launch")) return rc;
  }
  if (d.nsplit > d.nsplit_long) {
    const unsigned g = (unsigned)((d.nsplit - d.nsplit_long + kWG - 1) / kWG);
    hipLaunchKernelGGL((k_cvr_combine<VT, COMP>), dim3(g), dim3(kWG), 0, s, d.split_rows,
                       d.split_ptr, d.nsplit_long, d.nsplit, (const AC*)part, (IO*)y, (AC)alpha,
                       (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_cvr_combine launch")) return rc;
  }
  if (d.nempty > 0) {
    hipLaunchKernelGGL((k_scale_rows<VT>), dim3((unsigned)((d.nempty + kWG - 1) / kWG)), dim3(kWG),
                       0, s, d.empty_rows, d.nempty, (IO*)y, (AC)beta);
    return hip_check(hipGetLastError(), "k_scale_rows launch");
  }
  return RVVLA_OK;
}

int launch_cvr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_cvr<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_cvr<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_cvr<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_cvr<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
