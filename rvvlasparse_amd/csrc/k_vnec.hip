// k_vnec.hip -- VNEC SpMV for gfx950: merge-path partitions with the
// partition's non-empty columns of x staged in LDS.
//
// Semantics follow the reference VNEC kernels (reference
// src/rvv/spmv_mtx_vnec.cpp:192-586): each merge-path partition gathers the
// x values of the columns it touches into a compact buffer (there one per
// thread, here in LDS), multiplies its entries through the compacted column
// ranks (16-bit here), finishes the rows that end inside it, and hands the
// partial sum of the row still open at its end on as a carry (reference
// :357-361, :441-445, serially added).  The D/S/L variants of the reference
// are lane shapes of a CPU SIMD unit; on gfx950 all three run this kernel.
// Rows crossing partitions are finished by one owner thread that adds the
// pieces in partition order (deterministic, no atomics):
// y = alpha*sum + beta*y.  A workgroup issues all of its partition's loads
// (column list, entries, row bounds) before the x gathers; with 32-bit
// accumulation each XCD walks a contiguous run of partitions.
#include "kcommon.h"

namespace rvvla {

static_assert(vnec_items(4) % kWG == 0 && vnec_items(8) % kWG == 0,
              "k_vnec stages a whole number of columns and entries per thread");

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_vnec(
    const int32_t* __restrict__ cx, const int32_t* __restrict__ cy,
    const int32_t* __restrict__ nec_ptr, const int32_t* __restrict__ use_x,
    const uint16_t* __restrict__ ecr, const int32_t* __restrict__ rst,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ x,
    typename Val<VT>::io* __restrict__ y, typename Val<VT>::acc* __restrict__ first_part,
    typename Val<VT>::acc* __restrict__ fwd_part, int64_t nparts, int64_t m, typename Val<VT>::acc alpha,
    typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  constexpr int kItems = vnec_items((int)sizeof(A));
  __shared__ A xs[kItems];
  __shared__ A prod[kItems];
  const int t = sizeof(A) < 8 ? xcd_run_index((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int tid = threadIdx.x;
  if (t >= nparts) return;
  const int x0 = cx[t], x1 = cx[t + 1], y0 = cy[t], y1 = cy[t + 1];
  const int k0 = nec_ptr[t], nec = nec_ptr[t + 1] - k0;
  constexpr int KU = kItems / kWG;
  // every load of the partition in flight together: its column list, its
  // entries, the thread's row bounds; then the x gathers
  int u[KU];
#pragma unroll
  for (int j = 0; j < KU; ++j) {
    const int k = tid + j * kWG;
    u[j] = k < nec ? use_x[k0 + k] : -1;
  }
  const int n = y1 - y0;  // <= kItems
  A v[KU];
  uint16_t c[KU];
#pragma unroll
  for (int j = 0; j < KU; ++j) {
    const int e = y0 + tid + j * kWG;
    c[j] = e < y1 ? ecr[e] : (uint16_t)0;
    v[j] = e < y1 ? Val<VT>::get(val, e) : A(0);
  }
  const int nrows = x1 - x0 + (x1 < m ? 1 : 0);
  int lo = 0, hi = 0;
  if (nrows > kWG / 2 && tid < nrows) {
    const int r = x0 + tid;
    lo = max(rst[r], y0) - y0;
    hi = min(rst[r + 1], y1) - y0;
  }
  A gx[KU];
#pragma unroll
  for (int j = 0; j < KU; ++j) gx[j] = u[j] >= 0 ? (A)x[u[j]] : A(0);
#pragma unroll
  for (int j = 0; j < KU; ++j)
    if (u[j] >= 0) xs[tid + j * kWG] = gx[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < KU; ++j)
    if (tid + j * kWG < n) prod[tid + j * kWG] = v[j] * xs[c[j]];
  __syncthreads();
  auto finish = [&](int r, A s) {
    if (r == x1) {
      fwd_part[t] = s;  // the row still open at the partition's end
    } else if (r == x0 && rst[r] < y0) {
      first_part[t] = s;  // the row continued from earlier partitions
    } else {
      y[r] = alpha * s + scaled_y(beta, y, r);
    }
  };
  if (nrows > kWG / 2) {
    // a partition may hold up to kItems (empty) rows: stride over them
    for (int i = tid; i < nrows; i += kWG) {
      const int r = x0 + i;
      const int l = i == tid ? lo : max(rst[r], y0) - y0;
      const int h = i == tid ? hi : min(rst[r + 1], y1) - y0;
      Acc<A, COMP> s;
      for (int k = l; k < h; ++k) s.add(prod[k]);
      finish(r, s.value());
    }
    return;
  }
  if (nrows == 0) return;
  int width = 64;
  while (width * nrows > kWG) width >>= 1;
  const int g = tid / width, l = tid & (width - 1);
  Acc<A, COMP> a;
  if (g < nrows) {
    const int r = x0 + g;
    const int lo = max(rst[r], y0) - y0, hi = min(rst[r + 1], y1) - y0;
    for (int k = lo + l; k < hi; k += width) a.add(prod[k]);
  }
  a = group_reduce(a, width);
  if (g < nrows && l == 0) finish(x0 + g, a.value());
}

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_vnec_fixup(
    const int32_t* __restrict__ cx, const int32_t* __restrict__ cy,
    const int32_t* __restrict__ rst, typename Val<VT>::io* __restrict__ y,
    const typename Val<VT>::acc* __restrict__ first_part,
    const typename Val<VT>::acc* __restrict__ fwd_part, int64_t nparts, int64_t m,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t >= nparts) return;
  const int r = cx[t + 1];
  if (r >= m || rst[r] < cy[t] || rst[r] >= cy[t + 1]) return;  // not the row's first partition
  Acc<A, COMP> s;
  s.add(fwd_part[t]);
  for (int64_t q = t + 1; q < nparts; ++q) {
    if (cx[q] == cx[q + 1]) {
      s.add(fwd_part[q]);  // the whole partition lies inside the row
    } else {
      s.add(first_part[q]);
      break;
    }
  }
  y[r] = alpha * s.value() + scaled_y(beta, y, r);
}

template <class VT>
static int run_vnec(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevVNEC& d = A.vnec;
  if (d.nparts == 0) return RVVLA_OK;
  constexpr bool COMP = sizeof(VT) == 2;  // 16-bit values: TwoSum throughout
  AC* fp = (AC*)(scratch ? scratch : d.scratch);
  AC* wp = fp + d.nparts;
  hipLaunchKernelGGL((k_vnec<VT, COMP>), dim3((unsigned)(sizeof(AC) < 8 ? xcd_grid(d.nparts) : d.nparts)), dim3(kWG), 0, s, d.cx, d.cy,
                     d.nec_ptr, d.use_x, d.ecr, d.rst, (const VT*)d.val, (const IO*)x, (IO*)y, fp,
                     wp, d.nparts, A.m, (AC)alpha, (AC)beta);
  if (int rc = hip_check(hipGetLastError(), "k_vnec launch")) return rc;
  hipLaunchKernelGGL((k_vnec_fixup<VT, COMP>), dim3((unsigned)((d.nparts + kWG - 1) / kWG)), dim3(kWG),
                     0, s, d.cx, d.cy, d.rst, (IO*)y, fp, wp, d.nparts, A.m, (AC)alpha, (AC)beta);
  return hip_check(hipGetLastError(), "k_vnec_fixup launch");
}

int launch_vnec(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_vnec<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_vnec<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_vnec<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_vnec<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
