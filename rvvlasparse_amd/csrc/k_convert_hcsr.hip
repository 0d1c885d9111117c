// k_convert_hcsr.hip -- device-side CSR -> HCSR builder.
//
// The host converter (convert_hcsr.cpp, the reference's crs_to_HCSR) walks
// every entry twice on the CPU, plus the upload's re-lay.  This builder does
// both on the device from a CSR already there, with sort, search and scan
// passes only:
//  1. the row of every entry (binary search in the row pointers) and its
//     block key br * b_n + bc; a stable radix sort of (key, entry) puts the
//     entries block by block, each block in source CSR order -- the order
//     in which the reference converter fills a block (crs_to_hcsr.cpp
//     :107-119);
//  2. block_ptr[b] = the first sorted entry with key >= b (binary search);
//     kind and row_ptr length per block by the reference rule (CSR kind when
//     rows * p < block nnz, in double, :70-76), an exclusive scan of the
//     lengths gives row_offset;
//  3. one thread per sorted entry writes col_idx, values, its local row and
//     (COO kind) its row_ptr slot; one thread per row_ptr slot of a CSR-kind
//     block writes the count of the block's entries in earlier local rows
//     (binary search over the block's local rows, which ascend);
//  4. the kernels' image is then built from that layout by the host image
//     builder (hcsr_image.cpp) after a copy back (rvvla_hcsr_from_device).
// The reference layout (block_ptr, row_offset, row_ptr, col_idx, values)
// is bit-identical to the host converter's for the same R, C, p.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

// first index i in [lo, hi) with a[i] >= v (hi if none); a ascending
template <class T, class K>
__device__ __forceinline__ int64_t dev_lower_bound(const T* __restrict__ a, int64_t lo, int64_t hi,
                                                   K v) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if ((K)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_hc_keys(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                          int64_t m, int64_t nnz, int64_t R, int64_t C, int64_t bn,
                          int32_t* __restrict__ erow, uint32_t* __restrict__ key,
                          uint32_t* __restrict__ idx) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e >= nnz) return;
  // the row of entry e: the last r with rst[r] <= e (rst[0] = 0 <= e < rst[m])
  int64_t lo = 0, hi = m;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rst[mid] <= e) lo = mid;
    else hi = mid;
  }
  erow[e] = (int32_t)lo;
  key[e] = (uint32_t)((lo / R) * bn + col[e] / C);
  idx[e] = (uint32_t)e;
}

// block_ptr, kind and row_ptr length of block b (b = nblocks: the sentinel)
__global__ void k_hc_blocks(const uint32_t* __restrict__ skey, int64_t nnz, int64_t nblocks,
                            int64_t bn, int64_t m, int64_t R, double p,
                            int32_t* __restrict__ bp, uint8_t* __restrict__ kind,
                            int64_t* __restrict__ len) {
  const int64_t b = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (b > nblocks) return;
  const int64_t first = dev_lower_bound(skey, 0, nnz, (uint64_t)b);
  bp[b] = (int32_t)first;
  if (b == nblocks) return;
  const int64_t c = dev_lower_bound(skey, first, nnz, (uint64_t)b + 1) - first;
  const int64_t br = b / bn;
  const int64_t rows = (m - br * R) < R ? (m - br * R) : R;
  uint8_t k = 0;
  int64_t l = 0;
  if (c != 0) {
    if ((double)rows * p < (double)c) {
      k = 1;
      l = rows + 1;
    } else {
      k = 2;
      l = c;
    }
  }
  kind[b] = k;
  len[b] = l;
}

__global__ void k_hc_row_offset(const uint8_t* __restrict__ kind, const int64_t* __restrict__ off,
                                int64_t nblocks, int64_t total, uint32_t* __restrict__ ro) {
  const int64_t b = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (b > nblocks) return;
  ro[b] = b < nblocks ? ((uint32_t)kind[b] << 30) | (uint32_t)off[b] : (uint32_t)total;
}

// sorted entry i: col_idx, value, local row and (COO kind) its row_ptr slot
template <class V, class S, int NARROW>
__global__ void k_hc_scatter(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                             const int32_t* __restrict__ erow, const int32_t* __restrict__ col,
                             const V* __restrict__ val, const int32_t* __restrict__ bp,
                             const uint8_t* __restrict__ kind, const int64_t* __restrict__ off,
                             int64_t nnz, int64_t bn, int64_t R, int64_t C,
                             uint16_t* __restrict__ col16, S* __restrict__ sval,
                             uint16_t* __restrict__ lrow,
                             uint32_t* __restrict__ row_ptr) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i >= nnz) return;
  const uint32_t src = sidx[i];
  const int64_t b = skey[i], br = b / bn, bc = b % bn;
  col16[i] = (uint16_t)(col[src] - bc * C);
  const V v = val[src];
  if constexpr (NARROW == 1) sval[i] = dev_bf16((float)v);
  else if constexpr (NARROW == 2) sval[i] = dev_f16((float)v);
  else sval[i] = v;
  const uint32_t lr = (uint32_t)(erow[src] - br * R);
  lrow[i] = (uint16_t)lr;
  if (kind[b] == 2) row_ptr[off[b] + (i - bp[b])] = lr;
}

// row_ptr slot q of a CSR-kind block: the block's entries in local rows
// below q - off[b] (the block owning slot q: the last b with off[b] <= q)
__global__ void k_hc_csr_rowptr(const int32_t* __restrict__ bp, const uint8_t* __restrict__ kind,
                                const int64_t* __restrict__ off, int64_t nblocks, int64_t total,
                                const uint16_t* __restrict__ lrow, uint32_t* __restrict__ row_ptr) {
  const int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= total) return;
  const int64_t b = dev_lower_bound(off, 0, nblocks, q + 1) - 1;
  if (kind[b] != 1) return;
  const int64_t lr = q - off[b];
  row_ptr[q] = (uint32_t)(dev_lower_bound(lrow, bp[b], bp[b + 1], (uint64_t)lr) - bp[b]);
}

namespace {

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

// exclusive scan of cnt[0..n) into out; returns the total (out[n-1] + cnt[n-1])
int64_t scan_total(rvvla_mat_s& M, const int64_t* cnt, int64_t* out, int64_t n, hipStream_t s) {
  if (n == 0) return 0;
  size_t tmp = 0;
  int rc = hip_check(rocprim::exclusive_scan(nullptr, tmp, cnt, out, (int64_t)0, (size_t)n,
                                             rocprim::plus<int64_t>(), s),
                     "scan size");
  if (rc) throw Error(rc, last_error());
  void* tbuf = M.temp(tmp);
  rc = hip_check(rocprim::exclusive_scan(tbuf, tmp, cnt, out, (int64_t)0, (size_t)n,
                                         rocprim::plus<int64_t>(), s),
                 "scan");
  if (rc) throw Error(rc, last_error());
  int64_t last[2] = {0, 0};
  rc = hip_check(hipMemcpyAsync(&last[0], out + n - 1, 8, hipMemcpyDeviceToHost, s), "total");
  if (!rc) rc = hip_check(hipMemcpyAsync(&last[1], cnt + n - 1, 8, hipMemcpyDeviceToHost, s), "total");
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "total");
  if (rc) throw Error(rc, last_error());
  return last[0] + last[1];
}

void launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

}  // namespace

template <class V>
int build_hcsr_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t R, int64_t C, double p,
                      hipStream_t s) {
  DevHCSR& d = M.hcsr;
  const int64_t bm = (m + R - 1) / R, bn = (n + C - 1) / C, nblocks = bm * bn;
  d.R = R;
  d.C = C;
  d.b_m = bm;
  d.b_n = bn;
  // 1. rows, block keys, stable sort by block
  int32_t* erow = (int32_t*)M.temp((size_t)nnz * 4);
  uint32_t* key = (uint32_t*)M.temp((size_t)nnz * 16);
  uint32_t* idx = key + nnz;
  uint32_t* skey = idx + nnz;
  uint32_t* sidx = skey + nnz;
  if (nnz > 0) {
    hipLaunchKernelGGL(k_hc_keys, dim3(grid_of(nnz)), dim3(kWG), 0, s, rst, col, m, nnz, R, C, bn,
                       erow, key, idx);
    launched("k_hc_keys launch");
    unsigned end_bit = 1;
    while (end_bit < 32 && (int64_t(1) << end_bit) < nblocks) ++end_bit;
    size_t tmp = 0;
    int rc = hip_check(rocprim::radix_sort_pairs(nullptr, tmp, key, skey, idx, sidx, (size_t)nnz,
                                                 0, end_bit, s),
                       "radix sort size");
    if (rc) return rc;
    void* tbuf = M.temp(tmp);
    rc = hip_check(rocprim::radix_sort_pairs(tbuf, tmp, key, skey, idx, sidx, (size_t)nnz, 0,
                                             end_bit, s),
                   "radix sort");
    if (rc) return rc;
  }
  // 2. block_ptr, kinds, row_offset
  int32_t* bp = (int32_t*)M.put(nullptr, (size_t)(nblocks + 1) * 4);
  uint32_t* ro = (uint32_t*)M.put(nullptr, (size_t)(nblocks + 1) * 4);
  uint8_t* kind = (uint8_t*)M.temp((size_t)nblocks + 1);
  int64_t* len = (int64_t*)M.temp((size_t)(nblocks + 1) * 16);
  int64_t* off = len + nblocks + 1;
  hipLaunchKernelGGL(k_hc_blocks, dim3(grid_of(nblocks + 1)), dim3(kWG), 0, s, skey, nnz, nblocks,
                     bn, m, R, p, bp, kind, len);
  launched("k_hc_blocks launch");
  const int64_t total = scan_total(M, len, off, nblocks, s);
  if (total > ((int64_t(1) << 30) - 1)) {
    set_last_error("HCSR: row_ptr has " + std::to_string(total) +
                   " entries, more than the 30-bit row_offset field holds");
    return RVVLA_ERR_LENGTH;
  }
  hipLaunchKernelGGL(k_hc_row_offset, dim3(grid_of(nblocks + 1)), dim3(kWG), 0, s, kind, off,
                     nblocks, total, ro);
  launched("k_hc_row_offset launch");
  // 3. the entries
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint16_t* col16 = (uint16_t*)M.put(nullptr, (size_t)nnz * 2);
  void* sval = M.put(nullptr, (size_t)nnz * vbytes);
  uint16_t* lrow = (uint16_t*)M.temp((size_t)nnz * 2 + 2);
  uint32_t* row_ptr = (uint32_t*)M.put(nullptr, (size_t)total * 4);
  if (nnz > 0) {
    const unsigned g = grid_of(nnz);
    if (M.dtype == RVVLA_BF16)
      hipLaunchKernelGGL((k_hc_scatter<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, skey, sidx,
                         erow, col, val, bp, kind, off, nnz, bn, R, C, col16, (uint16_t*)sval,
                         lrow, row_ptr);
    else if (M.dtype == RVVLA_F16)
      hipLaunchKernelGGL((k_hc_scatter<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, skey, sidx,
                         erow, col, val, bp, kind, off, nnz, bn, R, C, col16, (uint16_t*)sval,
                         lrow, row_ptr);
    else
      hipLaunchKernelGGL((k_hc_scatter<V, V, 0>), dim3(g), dim3(kWG), 0, s, skey, sidx, erow, col,
                         val, bp, kind, off, nnz, bn, R, C, col16, (V*)sval, lrow, row_ptr);
    launched("k_hc_scatter launch");
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_hc_csr_rowptr, dim3(grid_of(total)), dim3(kWG), 0, s, bp, kind, off,
                       nblocks, total, lrow, row_ptr);
    launched("k_hc_csr_rowptr launch");
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device HCSR build")) return rc;
  M.free_temps();
  d.ref_block_ptr = bp;
  d.ref_row_offset = ro;
  d.ref_row_ptr = row_ptr;
  d.nref_row_ptr = total;
  d.ref_col_idx = col16;
  d.ref_values = sval;
  return RVVLA_OK;
}

template int build_hcsr_device<double>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int64_t, int64_t, double,
                                       hipStream_t);
template int build_hcsr_device<float>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int64_t, int64_t, double,
                                      hipStream_t);

}  // namespace rvvla
