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
//  4. the kernels' image from that layout: the non-empty block list and its
//     per-block-row offsets (a scan of non-empty flags), the entry-by-entry
//     block-rows (more than kHcsrFlat non-empty blocks) and the 256-row
//     units of the others, with the entry offset of every unit boundary in
//     their COO-kind blocks (binary search again).
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

// sorted entry i: col_idx, value, local row, (COO kind) its row_ptr slot,
// and (when some block-row is walked entry by entry) its block column
template <class V, class S, int NARROW>
__global__ void k_hc_scatter(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                             const int32_t* __restrict__ erow, const int32_t* __restrict__ col,
                             const V* __restrict__ val, const int32_t* __restrict__ bp,
                             const uint8_t* __restrict__ kind, const int64_t* __restrict__ off,
                             int64_t nnz, int64_t bn, int64_t R, int64_t C,
                             uint16_t* __restrict__ col16, S* __restrict__ sval,
                             uint16_t* __restrict__ lrow, uint16_t* __restrict__ ebc16,
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
  if (ebc16) ebc16[i] = (uint16_t)bc;
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

__global__ void k_hc_nonempty(const int32_t* __restrict__ bp, int64_t nblocks,
                              int64_t* __restrict__ flag) {
  const int64_t b = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (b < nblocks) flag[b] = bp[b + 1] > bp[b] ? 1 : 0;
}

// block-row offsets into the non-empty block list, and which block-rows
// are walked entry by entry (more than kHcsrFlat blocks, b_n <= 65536)
__global__ void k_hc_brow(const int64_t* __restrict__ nbidx, const int64_t* __restrict__ flag,
                          int64_t bm, int64_t bn, int64_t nblk, int64_t m, int64_t R,
                          int32_t* __restrict__ brow_ptr, uint8_t* __restrict__ flat,
                          int64_t* __restrict__ ucnt, int64_t* __restrict__ fcnt) {
  const int64_t br = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (br > bm) return;
  const int64_t first = br == bm ? nblk : bn > 0 ? nbidx[br * bn] : 0;
  brow_ptr[br] = (int32_t)first;
  if (br == bm) return;
  const int64_t last = bn > 0 ? nbidx[br * bn + bn - 1] + flag[br * bn + bn - 1] : first;
  const bool f = bn <= 65536 && last - first > kHcsrFlat;
  flat[br] = f ? 1 : 0;
  const int64_t rows = (m - br * R) < R ? (m - br * R) : R;
  ucnt[br] = f ? 0 : (rows + kWG - 1) / kWG;
  fcnt[br] = f ? 1 : 0;
}

// the non-empty block list; COO-kind blocks of unit block-rows ask for
// S + 1 unit-boundary offsets
__global__ void k_hc_list(const int32_t* __restrict__ bp, const uint8_t* __restrict__ kind,
                          const int64_t* __restrict__ off, const int64_t* __restrict__ nbidx,
                          const uint8_t* __restrict__ flat, int64_t nblocks, int64_t bn,
                          int64_t S, int32_t* __restrict__ nb_col, int32_t* __restrict__ nb_ptr,
                          uint32_t* __restrict__ nb_roff, int64_t* __restrict__ ccnt) {
  const int64_t b = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (b >= nblocks || bp[b + 1] == bp[b]) return;
  const int64_t k = nbidx[b];
  nb_col[k] = (int32_t)(b % bn);
  nb_ptr[k] = bp[b];
  const bool coo_unit = kind[b] == 2 && !flat[b / bn];
  nb_roff[k] = kind[b] == 1 ? ((1u << 30) | (uint32_t)off[b]) : (2u << 30);
  ccnt[k] = coo_unit ? S + 1 : 0;
}

__global__ void k_hc_csplit(const int32_t* __restrict__ bp, const uint8_t* __restrict__ kind,
                            const int64_t* __restrict__ nbidx, const uint8_t* __restrict__ flat,
                            const int64_t* __restrict__ coff, const uint16_t* __restrict__ lrow,
                            int64_t nblocks, int64_t bn, int64_t S, uint32_t* __restrict__ nb_roff,
                            int32_t* __restrict__ csplit) {
  const int64_t b = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (b >= nblocks || bp[b + 1] == bp[b] || kind[b] != 2 || flat[b / bn]) return;
  const int64_t k = nbidx[b], o = coff[k];
  nb_roff[k] = (2u << 30) | (uint32_t)o;
  int64_t e = bp[b];
  for (int64_t s = 0; s <= S; ++s) {
    e = dev_lower_bound(lrow, e, (int64_t)bp[b + 1], (uint64_t)(s * kWG));
    csplit[o + s] = (int32_t)(e - bp[b]);
  }
}

__global__ void k_hc_units(const int64_t* __restrict__ ucnt, const int64_t* __restrict__ ubase,
                           const int64_t* __restrict__ fcnt, const int64_t* __restrict__ fbase,
                           int64_t bm, int32_t* __restrict__ u_br, int32_t* __restrict__ u_lr0,
                           int32_t* __restrict__ flat_rows) {
  const int64_t br = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (br >= bm) return;
  for (int64_t u = 0; u < ucnt[br]; ++u) {
    u_br[ubase[br] + u] = (int32_t)br;
    u_lr0[ubase[br] + u] = (int32_t)(u * kWG);
  }
  if (fcnt[br]) flat_rows[fbase[br]] = (int32_t)br;
}

__global__ void k_hc_sentinel(int32_t* __restrict__ nb_ptr, int64_t nblk, int64_t nnz) {
  if (threadIdx.x == 0) nb_ptr[nblk] = (int32_t)nnz;
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
  const int64_t S = (R + kWG - 1) / kWG;
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
  // the non-empty blocks, the entry-by-entry block-rows, the units
  int64_t* flag = (int64_t*)M.temp((size_t)nblocks * 16 + 8);
  int64_t* nbidx = flag + nblocks;
  hipLaunchKernelGGL(k_hc_nonempty, dim3(grid_of(nblocks)), dim3(kWG), 0, s, bp, nblocks, flag);
  launched("k_hc_nonempty launch");
  const int64_t nblk = scan_total(M, flag, nbidx, nblocks, s);
  int32_t* brow_ptr = (int32_t*)M.put(nullptr, (size_t)(bm + 1) * 4);
  uint8_t* flat = (uint8_t*)M.temp((size_t)bm + 1);
  int64_t* bcnt = (int64_t*)M.temp((size_t)bm * 32 + 8);
  int64_t *ucnt = bcnt, *ubase = bcnt + bm, *fcnt = bcnt + 2 * bm, *fbase = bcnt + 3 * bm;
  hipLaunchKernelGGL(k_hc_brow, dim3(grid_of(bm + 1)), dim3(kWG), 0, s, nbidx, flag, bm, bn, nblk,
                     m, R, brow_ptr, flat, ucnt, fcnt);
  launched("k_hc_brow launch");
  const int64_t nunits = scan_total(M, ucnt, ubase, bm, s);
  const int64_t nflat = scan_total(M, fcnt, fbase, bm, s);
  // 3. the entries
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint16_t* col16 = (uint16_t*)M.put(nullptr, (size_t)nnz * 2);
  void* sval = M.put(nullptr, (size_t)nnz * vbytes);
  uint16_t* lrow = (uint16_t*)M.put(nullptr, (size_t)nnz * 2);
  uint16_t* ebc16 = nflat > 0 ? (uint16_t*)M.put(nullptr, (size_t)nnz * 2) : nullptr;
  uint32_t* row_ptr = (uint32_t*)M.put(nullptr, (size_t)total * 4);
  if (nnz > 0) {
    const unsigned g = grid_of(nnz);
    if (M.dtype == RVVLA_BF16)
      hipLaunchKernelGGL((k_hc_scatter<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, skey, sidx,
                         erow, col, val, bp, kind, off, nnz, bn, R, C, col16, (uint16_t*)sval,
                         lrow, ebc16, row_ptr);
    else if (M.dtype == RVVLA_F16)
      hipLaunchKernelGGL((k_hc_scatter<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, skey, sidx,
                         erow, col, val, bp, kind, off, nnz, bn, R, C, col16, (uint16_t*)sval,
                         lrow, ebc16, row_ptr);
    else
      hipLaunchKernelGGL((k_hc_scatter<V, V, 0>), dim3(g), dim3(kWG), 0, s, skey, sidx, erow, col,
                         val, bp, kind, off, nnz, bn, R, C, col16, (V*)sval, lrow, ebc16, row_ptr);
    launched("k_hc_scatter launch");
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_hc_csr_rowptr, dim3(grid_of(total)), dim3(kWG), 0, s, bp, kind, off,
                       nblocks, total, lrow, row_ptr);
    launched("k_hc_csr_rowptr launch");
  }
  // 4. the block list, unit boundaries, units
  int32_t* nb_col = (int32_t*)M.put(nullptr, (size_t)nblk * 4);
  int32_t* nb_ptr = (int32_t*)M.put(nullptr, (size_t)(nblk + 1) * 4);
  uint32_t* nb_roff = (uint32_t*)M.put(nullptr, (size_t)nblk * 4);
  int64_t* ccnt = (int64_t*)M.temp((size_t)nblk * 16 + 8);
  int64_t* coff = ccnt + nblk;
  hipLaunchKernelGGL(k_hc_list, dim3(grid_of(nblocks)), dim3(kWG), 0, s, bp, kind, off, nbidx,
                     flat, nblocks, bn, S, nb_col, nb_ptr, nb_roff, ccnt);
  launched("k_hc_list launch");
  hipLaunchKernelGGL(k_hc_sentinel, dim3(1), dim3(64), 0, s, nb_ptr, nblk, nnz);
  launched("k_hc_sentinel launch");
  const int64_t ncsplit = scan_total(M, ccnt, coff, nblk, s);
  if (ncsplit > 0x3FFFFFFF) {
    set_last_error("HCSR: csplit overflow");
    return RVVLA_ERR_LENGTH;
  }
  int32_t* csplit = (int32_t*)M.put(nullptr, (size_t)ncsplit * 4);
  hipLaunchKernelGGL(k_hc_csplit, dim3(grid_of(nblocks)), dim3(kWG), 0, s, bp, kind, nbidx, flat,
                     coff, lrow, nblocks, bn, S, nb_roff, csplit);
  launched("k_hc_csplit launch");
  int32_t* u_br = (int32_t*)M.put(nullptr, (size_t)nunits * 4);
  int32_t* u_lr0 = (int32_t*)M.put(nullptr, (size_t)nunits * 4);
  int32_t* flat_rows = (int32_t*)M.put(nullptr, (size_t)nflat * 4);
  hipLaunchKernelGGL(k_hc_units, dim3(grid_of(bm)), dim3(kWG), 0, s, ucnt, ubase, fcnt, fbase, bm,
                     u_br, u_lr0, flat_rows);
  launched("k_hc_units launch");
  uint8_t* flat_keep = nullptr;
  if (nflat > 0) {
    flat_keep = (uint8_t*)M.put(nullptr, (size_t)bm);
    if (int rc = hip_check(hipMemcpyAsync(flat_keep, flat, (size_t)bm, hipMemcpyDeviceToDevice, s),
                           "copy"))
      return rc;
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device HCSR build")) return rc;
  M.free_temps();
  d.brow_ptr = brow_ptr;
  d.nb_col = nb_col;
  d.nb_ptr = nb_ptr;
  d.nb_roff = nb_roff;
  d.rp32 = row_ptr;
  d.coo16 = lrow;
  d.col16 = col16;
  d.val = sval;
  d.flat = flat_keep;
  d.ebc16 = ebc16;
  d.nflat = nflat;
  d.flat_rows = flat_rows;
  d.nunits = nunits;
  d.nblk = nblk;
  d.nrp32 = total;
  d.ncsplit = ncsplit;
  d.u_br = u_br;
  d.u_lr0 = u_lr0;
  d.csplit = csplit;
  d.ref_block_ptr = bp;
  d.ref_row_offset = ro;
  return RVVLA_OK;
}

template int build_hcsr_device<double>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int64_t, int64_t, double,
                                       hipStream_t);
template int build_hcsr_device<float>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int64_t, int64_t, double,
                                      hipStream_t);

}  // namespace rvvla
