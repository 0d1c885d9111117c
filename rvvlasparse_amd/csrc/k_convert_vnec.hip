// k_convert_vnec.hip -- device-side CSR -> VNEC builder.
//
// The same layout as the host builder (convert_vnec.cpp) + upload, from a
// CSR already on the device:
//  1. the merge-path coordinate of every partition boundary (the search of
//     reference src/rvv/spmv_mtx_vnec.cpp:42-71, one thread each);
//  2. one workgroup per partition (at most vnec_items entries, the LDS
//     staging limit of the product kernel): its columns sorted in LDS
//     (bitonic), the distinct ones ranked by a prefix count -- the
//     ascending touched-column list -- and every entry's rank in that list
//     (binary search in LDS), the non-empty-column compaction of reference
//     include/VNEC.h:104-140;
//  3. a scan of the list lengths gives nec_ptr; the lists are copied into
//     place.
// The IRD variant selector of the reference is not evaluated here: the
// product kernel is the same for its three variants.
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

constexpr int kVnMax = 2048;  // entries of one partition (vnec_items of any dtype)
static_assert(vnec_items(4) <= kVnMax && vnec_items(8) <= kVnMax, "VNEC staging");
static_assert(kVnMax == 8 * kWG, "k_vn_compact: eight entries per thread");

__global__ void k_vn_coords(const int32_t* __restrict__ rst, int64_t m, int64_t nnz,
                            int64_t nparts, int64_t per, int64_t items, int32_t* __restrict__ cx,
                            int32_t* __restrict__ cy) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > nparts) return;
  const int64_t diag = per * t < items ? per * t : items;
  // row_end[i] = rst[i + 1]
  int64_t lo = diag - nnz > 0 ? diag - nnz : 0, hi = diag < m ? diag : m;
  while (lo < hi) {
    const int64_t pivot = (lo + hi) >> 1;
    if ((int64_t)rst[pivot + 1] <= diag - pivot - 1) lo = pivot + 1;
    else hi = pivot;
  }
  cx[t] = (int32_t)(lo < m ? lo : m);
  cy[t] = (int32_t)(diag - lo);
}

// one workgroup per partition: the sorted distinct columns (into
// lists[cy[t] ..]), their count, and each entry's rank among them
__global__ __launch_bounds__(kWG) void k_vn_compact(const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ cy,
                                                    int32_t* __restrict__ lists,
                                                    int64_t* __restrict__ nec,
                                                    uint16_t* __restrict__ ecr) {
  __shared__ int32_t keys[kVnMax];
  __shared__ int32_t uniq[kVnMax];
  __shared__ int32_t wtot[kWG / 64];
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t y0 = cy[t];
  const int cnt = (int)(cy[t + 1] - y0);  // <= kVnMax (checked on the host)
  int N = 1;
  while (N < cnt) N <<= 1;
  for (int i = tid; i < N; i += kWG) keys[i] = i < cnt ? col[y0 + i] : INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < N; i += kWG) {
        const int p = i ^ j;
        if (p > i) {
          const int32_t a = keys[i], b = keys[p];
          if ((a > b) == ((i & k) == 0)) {
            keys[i] = b;
            keys[p] = a;
          }
        }
      }
      __syncthreads();
    }
  // distinct values: a flag at the first of each run, ranked by a prefix
  // count (eight consecutive positions per thread)
  int f[8], c = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = tid * 8 + q;
    f[q] = i < cnt && (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
    c += f[q];
  }
  int incl = c;
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wv; ++w) before += wtot[w];
  int r = before + incl - c;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (f[q]) uniq[r++] = keys[tid * 8 + q];
  __syncthreads();
  const int total = wtot[0] + wtot[1] + wtot[2] + wtot[3];
  for (int i = tid; i < total; i += kWG) lists[y0 + i] = uniq[i];
  if (tid == 0) nec[t] = total;
  for (int i = tid; i < cnt; i += kWG) {
    const int32_t v = col[y0 + i];
    int lo = 0, hi = total;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (uniq[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    ecr[y0 + i] = (uint16_t)lo;
  }
}

__global__ void k_vn_place(const int32_t* __restrict__ lists, const int32_t* __restrict__ cy,
                           const int64_t* __restrict__ nec, const int64_t* __restrict__ base,
                           int64_t nparts, int32_t* __restrict__ nec_ptr,
                           int32_t* __restrict__ use_x) {
  const int64_t t = blockIdx.x;
  if (t > nparts) return;
  if (t == nparts) {
    if (threadIdx.x == 0) nec_ptr[t] = nparts > 0 ? (int32_t)(base[t - 1] + nec[t - 1]) : 0;
    return;
  }
  if (threadIdx.x == 0) nec_ptr[t] = (int32_t)base[t];
  for (int64_t i = threadIdx.x; i < nec[t]; i += kWG) use_x[base[t] + i] = lists[cy[t] + i];
}

template <class V, class S, int NARROW>
__global__ void k_vn_values(const V* __restrict__ val, int64_t nnz, S* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e >= nnz) return;
  const V v = val[e];
  if constexpr (NARROW == 1) out[e] = dev_bf16((float)v);
  else if constexpr (NARROW == 2) out[e] = dev_f16((float)v);
  else out[e] = v;
}

namespace {

unsigned vn_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

void vn_launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

}  // namespace

template <class V>
int build_vnec_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t nparts, hipStream_t s) {
  DevVNEC& d = M.vnec;
  const int64_t items = m + nnz, per = (items + nparts - 1) / nparts;
  const int64_t cap = vnec_items(M.dtype == RVVLA_F64 ? 8 : 4);
  if (per > cap) {
    set_last_error("VNEC: a partition exceeds the LDS staging size");
    return RVVLA_ERR_UNSUPPORTED_SPMV;
  }
  int32_t* cx = (int32_t*)M.put(nullptr, (size_t)(nparts + 1) * 4);
  int32_t* cy = (int32_t*)M.put(nullptr, (size_t)(nparts + 1) * 4);
  hipLaunchKernelGGL(k_vn_coords, dim3(vn_grid(nparts + 1)), dim3(kWG), 0, s, rst, m, nnz, nparts,
                     per, items, cx, cy);
  vn_launched("k_vn_coords launch");
  int32_t* lists = (int32_t*)M.temp((size_t)nnz * 4);
  int64_t* nec = (int64_t*)M.temp((size_t)nparts * 16 + 8);
  int64_t* base = nec + nparts;
  uint16_t* ecr = (uint16_t*)M.put(nullptr, (size_t)nnz * 2);
  hipLaunchKernelGGL(k_vn_compact, dim3((unsigned)nparts), dim3(kWG), 0, s, col, cy, lists, nec,
                     ecr);
  vn_launched("k_vn_compact launch");
  size_t tmp = 0;
  int rc = hip_check(rocprim::exclusive_scan(nullptr, tmp, nec, base, (int64_t)0, (size_t)nparts,
                                             rocprim::plus<int64_t>(), s),
                     "scan size");
  if (rc) return rc;
  void* tbuf = M.temp(tmp);
  rc = hip_check(rocprim::exclusive_scan(tbuf, tmp, nec, base, (int64_t)0, (size_t)nparts,
                                         rocprim::plus<int64_t>(), s),
                 "scan");
  if (rc) return rc;
  int64_t last[2] = {0, 0};
  rc = hip_check(hipMemcpyAsync(&last[0], base + nparts - 1, 8, hipMemcpyDeviceToHost, s), "total");
  if (!rc) rc = hip_check(hipMemcpyAsync(&last[1], nec + nparts - 1, 8, hipMemcpyDeviceToHost, s), "total");
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "total");
  if (rc) return rc;
  const int64_t nlist = last[0] + last[1];
  int32_t* nec_ptr = (int32_t*)M.put(nullptr, (size_t)(nparts + 1) * 4);
  int32_t* use_x = (int32_t*)M.put(nullptr, (size_t)nlist * 4);
  hipLaunchKernelGGL(k_vn_place, dim3((unsigned)(nparts + 1)), dim3(kWG), 0, s, lists, cy, nec,
                     base, nparts, nec_ptr, use_x);
  vn_launched("k_vn_place launch");
  int32_t* orst = (int32_t*)M.put(nullptr, (size_t)(m + 1) * 4);
  rc = hip_check(hipMemcpyAsync(orst, rst, (size_t)(m + 1) * 4, hipMemcpyDeviceToDevice, s), "copy");
  if (rc) return rc;
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  void* oval = M.put(nullptr, (size_t)nnz * vbytes);
  if (nnz > 0) {
    const unsigned g = vn_grid(nnz);
    if (M.dtype == RVVLA_BF16)
      hipLaunchKernelGGL((k_vn_values<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, val, nnz,
                         (uint16_t*)oval);
    else if (M.dtype == RVVLA_F16)
      hipLaunchKernelGGL((k_vn_values<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, val, nnz,
                         (uint16_t*)oval);
    else
      hipLaunchKernelGGL((k_vn_values<V, V, 0>), dim3(g), dim3(kWG), 0, s, val, nnz, (V*)oval);
    vn_launched("k_vn_values launch");
  }
  rc = hip_check(hipStreamSynchronize(s), "device VNEC build");
  if (rc) return rc;
  M.free_temps();
  const size_t acc = M.dtype == RVVLA_F64 ? 8 : 4;
  d.nparts = nparts;
  d.cx = cx;
  d.cy = cy;
  d.nec_ptr = nec_ptr;
  d.use_x = use_x;
  d.nlist = nlist;
  d.ecr = ecr;
  d.rst = orst;
  d.val = oval;
  d.scratch = M.put(nullptr, (size_t)nparts * 2 * acc + 16);
  M.device_bytes -= (int64_t)((size_t)nparts * 2 * acc + 16);
  return RVVLA_OK;
}

template int build_vnec_device<double>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int64_t, hipStream_t);
template int build_vnec_device<float>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int64_t, hipStream_t);

}  // namespace rvvla
