// k_convert.hip -- device-side CSR -> SELL-C-sigma (C = 64) builder.
//
// The host builder (convert_sell.cpp) takes about 0.3 s on a 4.2 M-row
// matrix, thousands of products' worth.  This one builds the same device
// image from a CSR that is already on the device, with a histogram-free
// pipeline of sort, scan and scatter:
//  1. key of row r = (window r / sigma) << 32 | (2^32 - 1 - len(r)); a
//     stable radix sort of (key, r) gives the permutation: rows of a sigma
//     window by descending length, ties in row order -- the host builder's
//     stable_sort by descending length within each window (sigma = 1 skips
//     the sort: identity);
//  2. one wavefront per 64-row slice: slice width = the longest of its rows,
//     height, first sorted row, and width * height slots;
//  3. an exclusive scan of the slot counts gives each slice's first slot;
//  4. one wavefront per slice scatters its rows: slot (j, k) at
//     Cs + j*Bs + k holds entry j of sorted row k, or value 0 and the row's
//     last column (column 0 for an empty row) past the row's end.
// The arrays are bit-identical to the host builder's for the same sigma;
// with sigma <= 0 the default rule is evaluated on the device: sigma = 1
// when 64-row slices in natural order pad by at most 1 % of the entries,
// else sigma = m when sorting all rows by length at least halves the padded
// slots, else 16384 (the host rule, same figures).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

__global__ void k_conv_keys(const int32_t* __restrict__ rst, int64_t m, int sigma,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ row) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r >= m) return;
  const uint32_t len = (uint32_t)(rst[r + 1] - rst[r]);
  key[r] = ((uint64_t)(r / sigma) << 32) | (uint64_t)(0xFFFFFFFFu - len);
  row[r] = (uint32_t)r;
}

// CSR validity on the device: flag |= 1 for a bad row pointer, 2 for a
// column out of range (the uploads check the same on the host)
__global__ void k_conv_check(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                             int64_t m, int64_t n, int64_t nnz, int* __restrict__ flag) {
  const int64_t i0 = (int64_t)blockIdx.x * kWG + threadIdx.x, step = (int64_t)gridDim.x * kWG;
  for (int64_t i = i0; i <= m; i += step) {
    const int32_t a = rst[i];
    if ((i == 0 && a != 0) || (i == m && a != nnz) || (i < m && rst[i + 1] < a) || a < 0 || a > nnz)
      atomicOr(flag, 1);
  }
  for (int64_t e = i0; e < nnz; e += step)
    if (col[e] < 0 || col[e] >= n) atomicOr(flag, 2);
}

int check_csr_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                     const int32_t* col, hipStream_t s) {
  int* dflag = (int*)M.temp(8);
  if (int rc = hip_check(hipMemsetAsync(dflag, 0, 4, s), "clear")) return rc;
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((m + 1 + kWG - 1) / kWG, 65536));
  hipLaunchKernelGGL(k_conv_check, dim3((unsigned)grid), dim3(kWG), 0, s, rst, col, m, n, nnz,
                     dflag);
  int flag = 0;
  if (int rc = hip_check(hipMemcpyAsync(&flag, dflag, 4, hipMemcpyDeviceToHost, s), "check"))
    return rc;
  if (int rc = hip_check(hipStreamSynchronize(s), "check")) return rc;
  if (flag) {
    set_last_error(flag & 1 ? "CSR: bad row pointers (device check)"
                            : "CSR: column out of range (device check)");
    return RVVLA_ERR_INVALID_ARG;
  }
  return RVVLA_OK;
}

__global__ void k_conv_iota(int64_t m, uint32_t* __restrict__ row) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m) row[r] = (uint32_t)r;
}

// natural-order padding of 64-row slices (for the default sigma)
__global__ void k_conv_padding(const int32_t* __restrict__ rst, int64_t m,
                               unsigned long long* __restrict__ padded) {
  const int64_t slice = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t r = slice * 64 + lane;
  if (slice * 64 >= m) return;
  const int len = r < m ? rst[r + 1] - rst[r] : 0;
  int w = len;
  for (int off = 32; off > 0; off >>= 1) w = max(w, __shfl_xor(w, off, 64));
  const int height = (int)min<int64_t>(64, m - slice * 64);
  if (lane == 0) atomicAdd(padded, (unsigned long long)w * (unsigned long long)height);
}

// lengths of all rows, as keys sorting to descending order
__global__ void k_conv_len_keys(const int32_t* __restrict__ rst, int64_t m,
                                uint32_t* __restrict__ key) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m) key[r] = 0xFFFFFFFFu - (uint32_t)(rst[r + 1] - rst[r]);
}

// slots of SELL-64 over rows sorted by descending length
__global__ void k_conv_sorted_slots(const uint32_t* __restrict__ key, int64_t m,
                                    unsigned long long* __restrict__ slots) {
  const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
  const int64_t r0 = k * 64;
  if (r0 >= m) return;
  const unsigned long long w = 0xFFFFFFFFu - key[r0];
  atomicAdd(slots, w * (unsigned long long)min<int64_t>(64, m - r0));
}

__global__ void k_conv_slices(const int32_t* __restrict__ rst, const uint32_t* __restrict__ perm,
                              int64_t m, int64_t nslices, int32_t* __restrict__ cl,
                              int32_t* __restrict__ bs, int32_t* __restrict__ sbs,
                              int64_t* __restrict__ slots) {
  const int64_t slice = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (slice >= nslices) return;
  const int64_t k = slice * 64 + lane;
  int len = 0;
  if (k < m) {
    const uint32_t r = perm[k];
    len = rst[r + 1] - rst[r];
  }
  int w = len;
  for (int off = 32; off > 0; off >>= 1) w = max(w, __shfl_xor(w, off, 64));
  if (lane == 0) {
    const int height = (int)min<int64_t>(64, m - slice * 64);
    cl[slice] = w;
    bs[slice] = height;
    sbs[slice] = (int32_t)(slice * 64);
    slots[slice] = (int64_t)w * height;
  }
}

__global__ void k_conv_narrow_cs(const int64_t* __restrict__ cs64, int64_t nslices,
                                 int32_t* __restrict__ cs) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < nslices) cs[i] = (int32_t)cs64[i];
}

__global__ void k_conv_not_identity(const uint32_t* __restrict__ perm, int64_t m,
                                    int* __restrict__ flag) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m && perm[r] != (uint32_t)r) atomicOr(flag, 1);
}

// slots of a slice written by one wavefront: slices wider than this are cut
// into column blocks of kScatterCols, each its own wavefront
constexpr int kScatterCols = 256;

__global__ void k_conv_col_blocks(const int32_t* __restrict__ cl, int64_t nslices,
                                  int64_t* __restrict__ nblk) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < nslices) nblk[i] = cl[i] > kScatterCols ? (cl[i] + kScatterCols - 1) / kScatterCols : 1;
}

// V: input value type (double for F64, float otherwise); S: stored type
// (V, or uint16_t bits for BF16 / F16 with NARROW = 1 / 2).  Wavefront wv
// writes column block wv - bbase[slice] of the slice whose blocks hold it
// (bbase: exclusive scan of the per-slice block counts).
template <class V, class S, int NARROW>
__global__ void k_conv_scatter(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                               const V* __restrict__ val, const uint32_t* __restrict__ perm,
                               const int32_t* __restrict__ cl, const int32_t* __restrict__ cs,
                               const int64_t* __restrict__ bbase, int64_t nwaves, int64_t m,
                               int64_t nslices, uint32_t* __restrict__ scol,
                               S* __restrict__ sval) {
  const int64_t wv = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wv >= nwaves) return;
  int64_t lo = 0, hi = nslices;  // the last slice with bbase <= wv
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (bbase[mid] <= wv) lo = mid;
    else hi = mid;
  }
  const int64_t slice = lo;
  const int64_t k = slice * 64 + lane;
  if (k >= m) return;
  const int height = (int)min<int64_t>(64, m - slice * 64);
  const int j0 = (int)(wv - bbase[slice]) * kScatterCols;
  const int w = cl[slice] > kScatterCols ? min(cl[slice], j0 + kScatterCols) : cl[slice];
  const int64_t base = cs[slice];
  const uint32_t r = perm[k];
  const int e0 = rst[r], l = rst[r + 1] - e0;
  const uint32_t pad_col = l > 0 ? (uint32_t)col[e0 + l - 1] : 0u;
  for (int j = j0; j < w; ++j) {
    const int64_t slot = base + (int64_t)j * height + lane;
    const bool in = j < l;
    scol[slot] = in ? (uint32_t)col[e0 + j] : pad_col;
    const V v = in ? val[e0 + j] : V(0);
    if constexpr (NARROW == 1) sval[slot] = dev_bf16((float)v);
    else if constexpr (NARROW == 2) sval[slot] = dev_f16((float)v);
    else sval[slot] = v;
  }
}

template <class V>
int build_sell_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int sigma, hipStream_t s) {
  DevSELL& d = M.sell;
  const int64_t nslices = (m + 63) / 64;
  const unsigned g_rows = (unsigned)((m + kWG - 1) / kWG);
  const unsigned g_slices = (unsigned)((nslices + kWG / 64 - 1) / (kWG / 64));
  if (m == 0) {
    d.nslices = 0;
    d.perm_identity = true;
    return RVVLA_OK;
  }
  // default sigma: natural-order padding, then the padding of a full sort
  if (sigma <= 0) {
    unsigned long long* dpad = (unsigned long long*)M.temp(16);
    if (int rc = hip_check(hipMemsetAsync(dpad, 0, 16, s), "clear")) return rc;
    hipLaunchKernelGGL(k_conv_padding, dim3(g_slices), dim3(kWG), 0, s, rst, m, dpad);
    unsigned long long padded[2] = {0, 0};
    if (int rc = hip_check(hipMemcpyAsync(padded, dpad, 8, hipMemcpyDeviceToHost, s), "padding"))
      return rc;
    if (int rc = hip_check(hipStreamSynchronize(s), "padding")) return rc;
    if (nnz > 0 && (double)((int64_t)padded[0] - nnz) <= 0.01 * (double)nnz) {
      sigma = 1;
    } else if (nnz == 0) {
      sigma = 16384;
    } else {
      uint32_t* key = (uint32_t*)M.temp(2 * (size_t)m * 4);
      uint32_t* key_out = key + m;
      hipLaunchKernelGGL(k_conv_len_keys, dim3(g_rows), dim3(kWG), 0, s, rst, m, key);
      size_t tmp = 0;
      if (int rc = hip_check(rocprim::radix_sort_keys(nullptr, tmp, key, key_out, (size_t)m, 0, 32, s),
                             "radix sort size"))
        return rc;
      void* tbuf = M.temp(tmp);
      if (int rc = hip_check(rocprim::radix_sort_keys(tbuf, tmp, key, key_out, (size_t)m, 0, 32, s),
                             "radix sort"))
        return rc;
      hipLaunchKernelGGL(k_conv_sorted_slots, dim3((unsigned)((nslices + kWG - 1) / kWG)), dim3(kWG), 0, s,
                         key_out, m, dpad + 1);
      if (int rc = hip_check(hipMemcpyAsync(&padded[1], dpad + 1, 8, hipMemcpyDeviceToHost, s), "padding"))
        return rc;
      if (int rc = hip_check(hipStreamSynchronize(s), "padding")) return rc;
      sigma = (double)padded[1] <= 0.5 * (double)padded[0] ? (int)m : 16384;
    }
  }
  uint32_t* perm = (uint32_t*)M.put(nullptr, (size_t)m * 4);
  if (sigma == 1) {
    hipLaunchKernelGGL(k_conv_iota, dim3(g_rows), dim3(kWG), 0, s, m, perm);
  } else {
    uint64_t* keys = (uint64_t*)M.temp(2 * (size_t)m * 8 + (size_t)m * 4);
    uint64_t* keys_out = keys + m;
    uint32_t* rows = (uint32_t*)(keys_out + m);
    hipLaunchKernelGGL(k_conv_keys, dim3(g_rows), dim3(kWG), 0, s, rst, m, sigma, keys, rows);
    const int64_t nwin = (m + sigma - 1) / sigma;
    unsigned end_bit = 32;
    while (end_bit < 64 && (int64_t(1) << (end_bit - 32)) < nwin) ++end_bit;
    size_t tmp = 0;
    if (int rc = hip_check(rocprim::radix_sort_pairs(nullptr, tmp, keys, keys_out, rows, perm,
                                                      (size_t)m, 0, end_bit, s),
                           "radix sort size"))
      return rc;
    void* tbuf = M.temp(tmp);
    if (int rc = hip_check(rocprim::radix_sort_pairs(tbuf, tmp, keys, keys_out, rows, perm,
                                                      (size_t)m, 0, end_bit, s),
                           "radix sort"))
      return rc;
  }
  int32_t* cl = (int32_t*)M.put(nullptr, (size_t)nslices * 4);
  int32_t* cs = (int32_t*)M.put(nullptr, (size_t)nslices * 4);
  int32_t* bs = (int32_t*)M.put(nullptr, (size_t)nslices * 4);
  int32_t* sbs = (int32_t*)M.put(nullptr, (size_t)nslices * 4);
  int64_t* slots = (int64_t*)M.temp(2 * (size_t)(nslices + 1) * 8);
  int64_t* cs64 = slots + nslices + 1;
  hipLaunchKernelGGL(k_conv_slices, dim3(g_slices), dim3(kWG), 0, s, rst, perm, m, nslices, cl,
                     bs, sbs, slots);
  size_t tmp = 0;
  if (int rc = hip_check(rocprim::exclusive_scan(nullptr, tmp, slots, cs64, (int64_t)0,
                                                  (size_t)nslices, rocprim::plus<int64_t>(), s),
                         "scan size"))
    return rc;
  void* tbuf = M.temp(tmp);
  if (int rc = hip_check(rocprim::exclusive_scan(tbuf, tmp, slots, cs64, (int64_t)0,
                                                  (size_t)nslices, rocprim::plus<int64_t>(), s),
                         "scan"))
    return rc;
  int64_t last[2] = {0, 0};
  if (int rc = hip_check(hipMemcpyAsync(&last[0], cs64 + nslices - 1, 8, hipMemcpyDeviceToHost, s), "total"))
    return rc;
  if (int rc = hip_check(hipMemcpyAsync(&last[1], slots + nslices - 1, 8, hipMemcpyDeviceToHost, s), "total"))
    return rc;
  if (int rc = hip_check(hipStreamSynchronize(s), "total")) return rc;
  const int64_t total = last[0] + last[1];
  if (total > 2147483647LL) {
    set_last_error("SELL: more than 2^31 slots");
    return RVVLA_ERR_LENGTH;
  }
  hipLaunchKernelGGL(k_conv_narrow_cs, dim3((unsigned)((nslices + kWG - 1) / kWG)), dim3(kWG), 0, s,
                     cs64, nslices, cs);
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint32_t* scol = (uint32_t*)M.put(nullptr, (size_t)total * 4);
  void* sval = M.put(nullptr, (size_t)total * vbytes);
  // column blocks of the wide slices, a wavefront each
  int64_t* nblk = (int64_t*)M.temp((size_t)nslices * 16);
  int64_t* bbase = nblk + nslices;
  hipLaunchKernelGGL(k_conv_col_blocks, dim3((unsigned)((nslices + kWG - 1) / kWG)), dim3(kWG), 0,
                     s, cl, nslices, nblk);
  tmp = 0;
  if (int rc = hip_check(rocprim::exclusive_scan(nullptr, tmp, nblk, bbase, (int64_t)0,
                                                  (size_t)nslices, rocprim::plus<int64_t>(), s),
                         "scan size"))
    return rc;
  tbuf = M.temp(tmp);
  if (int rc = hip_check(rocprim::exclusive_scan(tbuf, tmp, nblk, bbase, (int64_t)0,
                                                  (size_t)nslices, rocprim::plus<int64_t>(), s),
                         "scan"))
    return rc;
  if (int rc = hip_check(hipMemcpyAsync(&last[0], bbase + nslices - 1, 8, hipMemcpyDeviceToHost, s), "blocks"))
    return rc;
  if (int rc = hip_check(hipMemcpyAsync(&last[1], nblk + nslices - 1, 8, hipMemcpyDeviceToHost, s), "blocks"))
    return rc;
  if (int rc = hip_check(hipStreamSynchronize(s), "blocks")) return rc;
  const int64_t nwaves = last[0] + last[1];
  const unsigned g_waves = (unsigned)((nwaves + kWG / 64 - 1) / (kWG / 64));
  if (M.dtype == RVVLA_BF16)
    hipLaunchKernelGGL((k_conv_scatter<V, uint16_t, 1>), dim3(g_waves), dim3(kWG), 0, s, rst, col,
                       val, perm, cl, cs, bbase, nwaves, m, nslices, scol, (uint16_t*)sval);
  else if (M.dtype == RVVLA_F16)
    hipLaunchKernelGGL((k_conv_scatter<V, uint16_t, 2>), dim3(g_waves), dim3(kWG), 0, s, rst, col,
                       val, perm, cl, cs, bbase, nwaves, m, nslices, scol, (uint16_t*)sval);
  else
    hipLaunchKernelGGL((k_conv_scatter<V, V, 0>), dim3(g_waves), dim3(kWG), 0, s, rst, col, val,
                       perm, cl, cs, bbase, nwaves, m, nslices, scol, (V*)sval);
  int flag = 0;
  if (sigma > 1) {
    int* dflag = (int*)M.temp(8);
    if (int rc = hip_check(hipMemsetAsync(dflag, 0, 4, s), "clear")) return rc;
    hipLaunchKernelGGL(k_conv_not_identity, dim3(g_rows), dim3(kWG), 0, s, perm, m, dflag);
    if (int rc = hip_check(hipMemcpyAsync(&flag, dflag, 4, hipMemcpyDeviceToHost, s), "flag"))
      return rc;
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device SELL build")) return rc;
  M.free_temps();
  d.nslices = nslices;
  d.cl = cl;
  d.cs = cs;
  d.bs = bs;
  d.sbs = sbs;
  d.col = scol;
  d.perm = perm;
  d.val = sval;
  d.perm_identity = flag == 0;
  return RVVLA_OK;
}

template int build_sell_device<double>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int, hipStream_t);
template int build_sell_device<float>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int, hipStream_t);

}  // namespace rvvla
