// k_convert_csr5.hip -- device-side CSR -> CSR5 (omega = 64) builder.
//
// Builds on the device, from a CSR already there, the layout the host
// builder (convert_csr5.cpp: Liu & Vinter's method in the field layout the
// reference kernel reads, reference src/rvv/spmv_mtx_csr5.cpp:85-365) gives
// bit for bit, then the kernels' extras:
//  1. tile_ptr[t] = the row holding entry t * omega * sigma (binary search;
//     m past the end);
//  2. empty rows: a scan of empty-row flags; a full tile whose row range
//     has an empty row strictly inside gets bit 31 and as many empty-offset
//     slots as rows start strictly inside it (binary searches and the
//     empty-row prefix), and a scan of those counts gives the
//     empty-offset pointer;
//  3. one wavefront per full tile, lane l = CSR5 lane l: the row-start flags
//     of its sigma entries, its y_offset (a wavefront prefix sum of the
//     counted flags), its seg_offset (ballot of the lanes with counted
//     flags), the empty offsets of its flags, and the descriptor words;
//  4. the full tiles' entries transposed (entry i of lane l at
//     tile_base + i * omega + l), the tail copied;
//  5. the kernels' image: the empty-row list, and per tile where the
//     segment it leaves open ends (a suffix minimum over the tiles that
//     hold a row start) with the long segments and long tail rows listed
//     for a wavefront each.
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

constexpr int kC5Omega = 64;

// first index i in [lo, hi) with a[i] > v; a ascending
__device__ __forceinline__ int64_t c5_upper(const int32_t* __restrict__ a, int64_t lo, int64_t hi,
                                            int64_t v) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
// first index i in [lo, hi) with a[i] >= v
__device__ __forceinline__ int64_t c5_lower(const int32_t* __restrict__ a, int64_t lo, int64_t hi,
                                            int64_t v) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_c5_tile_ptr(const int32_t* __restrict__ rst, int64_t m, int64_t nnz, int64_t W,
                              int64_t p, uint32_t* __restrict__ tp) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > p) return;
  const int64_t e = t * W;
  tp[t] = e < nnz ? (uint32_t)(c5_upper(rst, 0, m + 1, e) - 1) : (uint32_t)m;
}

__global__ void k_c5_empty(const int32_t* __restrict__ rst, int64_t m, int64_t* __restrict__ ef) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m) ef[r] = rst[r] == rst[r + 1] ? 1 : 0;
}

// bit 31 of the full tiles with an empty row strictly inside their row
// range, and their empty-offset slot counts (rows starting strictly
// inside the tile that hold entries); tpf = tp with the flags
__global__ void k_c5_tile_empty(const int32_t* __restrict__ rst, const uint32_t* __restrict__ tp,
                                const int64_t* __restrict__ nempty, int64_t m, int64_t W,
                                int64_t ntiles, int64_t p, uint32_t* __restrict__ tpf,
                                int64_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > p) return;
  if (t >= ntiles) {
    tpf[t] = tp[t];
    return;
  }
  const int64_t r0 = tp[t], r1 = tp[t + 1];
  int64_t c = 0;
  uint32_t flag = 0;
  if (r1 > r0 + 1 && nempty[r1] - nempty[r0 + 1] > 0) {
    flag = 0x80000000u;
    const int64_t b = t * W;
    const int64_t lo = c5_upper(rst, 0, m + 1, b);
    int64_t hi = c5_lower(rst, 0, m + 1, b + W);
    if (hi > m) hi = m;
    if (hi > lo) c = (hi - lo) - (nempty[hi] - nempty[lo]);
  }
  tpf[t] = tp[t] | flag;
  cnt[t] = c;
}

__global__ void k_c5_eptr(const int64_t* __restrict__ escan, int64_t ntiles, int64_t total,
                          int64_t p, uint32_t* __restrict__ eptr) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > p) return;
  eptr[t] = t < ntiles ? (uint32_t)escan[t] : (t == ntiles && ntiles > 0) ? (uint32_t)total : 0u;
}

// one wavefront per full tile: flags, y_offset, seg_offset, empty offsets,
// descriptor words, and the tile's kind (bit 0: a row starts at its first
// entry, bit 1: a row starts later in it)
__global__ __launch_bounds__(kWG) void k_c5_desc(const int32_t* __restrict__ rst,
                                                 const uint32_t* __restrict__ tpf,
                                                 const uint32_t* __restrict__ eptr, int64_t m,
                                                 int64_t ntiles, int sigma, int dl, int bit_y,
                                                 int bit_seg, uint32_t* __restrict__ desc,
                                                 uint32_t* __restrict__ eoff,
                                                 uint8_t* __restrict__ kind) {
  const int64_t t = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (t >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int64_t W = (int64_t)kC5Omega * sigma, base = t * W;
  const bool has_empty = (tpf[t] >> 31) != 0;
  const int64_t row0 = tpf[t] & 0x7FFFFFFFu;
  const int64_t start = base + (int64_t)lane * sigma;
  // the flags of this lane: entry e starts a row when it is the first entry
  // of the (non-empty) row holding it
  int64_t r = c5_upper(rst, 0, m + 1, start) - 1;
  uint64_t fl = 0;
  for (int i = 0; i < sigma; ++i) {
    const int64_t e = start + i;
    while (rst[r + 1] <= e) ++r;
    if (rst[r] == e) fl |= uint64_t(1) << i;
  }
  const uint64_t counted = lane == 0 ? (fl & ~uint64_t(1)) : fl;
  const int c = __popcll(counted);
  int incl = c;
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  const uint32_t yoff = (uint32_t)(incl - c);
  const uint64_t anym = __ballot(c > 0);
  const uint64_t after = lane == 63 ? 0ull : (anym & (~0ull << (lane + 1)));
  const int next = after ? __ffsll((unsigned long long)after) - 1 : kC5Omega;
  const uint32_t seg = (uint32_t)(next - lane - 1);
  if (has_empty && c > 0) {  // the k-th counted flag of the lane starts row row0 + 1 + eoff
    uint32_t k = 0;
    int64_t rr = c5_upper(rst, 0, m + 1, start) - 1;
    for (int i = 0; i < sigma; ++i) {
      const int64_t e = start + i;
      while (rst[rr + 1] <= e) ++rr;
      if (!((counted >> i) & 1u)) continue;
      eoff[eptr[t] + yoff + k] = (uint32_t)(rr - row0 - 1);
      ++k;
    }
  }
  const int f0 = 32 - bit_y - bit_seg;
  uint32_t w[3] = {(yoff << (32 - bit_y)) | (seg << (32 - bit_y - bit_seg)), 0u, 0u};
  for (int i = 0; i < sigma; ++i) {
    if (!((fl >> i) & 1u)) continue;
    if (i < f0) w[0] |= 1u << (31 - (32 - f0) - i);
    else w[1 + (i - f0) / 32] |= 1u << (31 - (i - f0) % 32);
  }
  uint32_t* dp = desc + t * kC5Omega * dl + lane;
  for (int k = 0; k < dl; ++k) dp[k * kC5Omega] = w[k];
  const uint64_t laterm = __ballot(counted != 0);
  if (lane == 0) kind[t] = (uint8_t)((fl & 1u) | (laterm ? 2u : 0u));
}

// the entries: full tiles transposed, the tail in place
template <class V, class S, int NARROW>
__global__ void k_c5_transpose(const int32_t* __restrict__ col, const V* __restrict__ val,
                               int64_t nnz, int64_t ntiles, int sigma,
                               uint32_t* __restrict__ ocol, S* __restrict__ oval) {
  const int64_t j = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (j >= nnz) return;
  const int64_t W = (int64_t)kC5Omega * sigma;
  int64_t src = j;
  if (j < ntiles * W) {
    const int64_t t = j / W, rem = j - t * W, i = rem / kC5Omega, l = rem - i * kC5Omega;
    src = t * W + l * sigma + i;
  }
  ocol[j] = (uint32_t)col[src];
  const V v = val[src];
  if constexpr (NARROW == 1) oval[j] = dev_bf16((float)v);
  else if constexpr (NARROW == 2) oval[j] = dev_f16((float)v);
  else oval[j] = v;
}

__global__ void k_c5_empty_list(const int64_t* __restrict__ ef, const int64_t* __restrict__ nempty,
                                int64_t m, int32_t* __restrict__ list) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m && ef[r]) list[nempty[r]] = (int32_t)r;
}

// a[j] = index of tile ntiles-1-j if it holds a row start, else ntiles
__global__ void k_c5_rev_stops(const uint8_t* __restrict__ kind, int64_t ntiles,
                               int64_t* __restrict__ a) {
  const int64_t j = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (j >= ntiles) return;
  const int64_t t = ntiles - 1 - j;
  a[j] = kind[t] ? t : ntiles;
}

// chain[t] from the suffix minima (sm[j] = min over tiles >= ntiles-1-j);
// the tiles whose open segment needs a wavefront (lflag), and the tail
// rows that do (tflag)
__global__ void k_c5_chain(const uint8_t* __restrict__ kind, const int64_t* __restrict__ sm,
                           int64_t ntiles, int32_t* __restrict__ chain,
                           int64_t* __restrict__ lflag) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t >= ntiles) return;
  const int64_t next = t + 1 < ntiles ? sm[ntiles - 2 - t] : ntiles;
  const uint32_t end = next == ntiles ? ((uint32_t)ntiles | 0x80000000u)
                       : (kind[next] & 1u) ? (uint32_t)next : (uint32_t)next + 1;
  chain[t] = (int32_t)end;
  lflag[t] = kind[t] != 0 && ((end >> 31) || (int64_t)(end & 0x7FFFFFFFu) - t - 1 > kCsr5Serial);
}

__global__ void k_c5_tail_flags(const int32_t* __restrict__ rst, int64_t first_tail_row, int64_t m,
                                int64_t* __restrict__ tflag) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  const int64_t r = first_tail_row + i;
  if (r < m) tflag[i] = rst[r + 1] - rst[r] > kCsr5Serial ? 1 : 0;
}

__global__ void k_c5_long_list(const int64_t* __restrict__ lflag, const int64_t* __restrict__ lpos,
                               int64_t ntiles, const int64_t* __restrict__ tflag,
                               const int64_t* __restrict__ tpos, int64_t ntail, int64_t nl,
                               int64_t first_tail_row, int32_t* __restrict__ longs) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < ntiles) {
    if (lflag[i]) longs[lpos[i]] = (int32_t)i;
  } else if (i < ntiles + ntail) {
    const int64_t j = i - ntiles;
    if (tflag[j]) longs[nl + tpos[j]] = (int32_t)(-(first_tail_row + j + 1));
  }
}

namespace {

unsigned c5_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

void c5_launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

int64_t c5_scan(rvvla_mat_s& M, const int64_t* cnt, int64_t* out, int64_t n, hipStream_t s) {
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

int c5_ceil_log2(int64_t v) {
  int b = 0;
  while ((int64_t(1) << b) < v) ++b;
  return b;
}

}  // namespace

template <class V>
int build_csr5_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int sigma, hipStream_t s) {
  DevCSR5& d = M.csr5;
  const int64_t W = (int64_t)kC5Omega * sigma;
  const int64_t p = nnz > 0 ? (nnz + W - 1) / W : 1, ntiles = p - 1;
  d.ntiles = ntiles;
  d.sigma = sigma;
  d.bit_y = c5_ceil_log2(W);
  d.bit_seg = c5_ceil_log2(kC5Omega);
  d.dl = (d.bit_y + d.bit_seg + sigma + 31) / 32;
  d.tail_base = ntiles * W;
  d.np = p;
  // 1-2. tile pointers, empty rows, empty-offset slots
  uint32_t* tp = (uint32_t*)M.temp((size_t)(p + 1) * 4);
  uint32_t* tpf = (uint32_t*)M.put(nullptr, (size_t)(p + 1) * 4);
  hipLaunchKernelGGL(k_c5_tile_ptr, dim3(c5_grid(p + 1)), dim3(kWG), 0, s, rst, m, nnz, W, p, tp);
  c5_launched("k_c5_tile_ptr launch");
  int64_t* ef = (int64_t*)M.temp((size_t)(2 * m + 1) * 8 + 8);
  int64_t* nempty = ef + m;
  hipLaunchKernelGGL(k_c5_empty, dim3(c5_grid(m)), dim3(kWG), 0, s, rst, m, ef);
  c5_launched("k_c5_empty launch");
  const int64_t total_empty = c5_scan(M, ef, nempty, m, s);
  if (int rc = hip_check(hipMemcpyAsync(nempty + m, &total_empty, 8, hipMemcpyHostToDevice, s),
                         "copy"))
    return rc;
  int64_t* cnt = (int64_t*)M.temp((size_t)(2 * ntiles + 1) * 8 + 8);
  int64_t* escan = cnt + ntiles;
  hipLaunchKernelGGL(k_c5_tile_empty, dim3(c5_grid(p + 1)), dim3(kWG), 0, s, rst, tp, nempty, m, W,
                     ntiles, p, tpf, cnt);
  c5_launched("k_c5_tile_empty launch");
  const int64_t noff = c5_scan(M, cnt, escan, ntiles, s);
  uint32_t* eptr = (uint32_t*)M.put(nullptr, (size_t)(p + 1) * 4);
  hipLaunchKernelGGL(k_c5_eptr, dim3(c5_grid(p + 1)), dim3(kWG), 0, s, escan, ntiles, noff, p, eptr);
  c5_launched("k_c5_eptr launch");
  // 3. descriptors and empty offsets
  uint32_t* desc = (uint32_t*)M.put(nullptr, (size_t)ntiles * kC5Omega * d.dl * 4);
  uint32_t* eoff = (uint32_t*)M.put(nullptr, (size_t)noff * 4);
  uint8_t* kind = (uint8_t*)M.temp((size_t)ntiles + 1);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_c5_desc, dim3((unsigned)((ntiles + kWG / 64 - 1) / (kWG / 64))),
                       dim3(kWG), 0, s, rst, tpf, eptr, m, ntiles, sigma, d.dl, d.bit_y, d.bit_seg,
                       desc, eoff, kind);
    c5_launched("k_c5_desc launch");
  }
  // 4. entries
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint32_t* ocol = (uint32_t*)M.put(nullptr, (size_t)nnz * 4);
  void* oval = M.put(nullptr, (size_t)nnz * vbytes);
  if (nnz > 0) {
    const unsigned g = c5_grid(nnz);
    if (M.dtype == RVVLA_BF16)
      hipLaunchKernelGGL((k_c5_transpose<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, col, val, nnz,
                         ntiles, sigma, ocol, (uint16_t*)oval);
    else if (M.dtype == RVVLA_F16)
      hipLaunchKernelGGL((k_c5_transpose<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, col, val, nnz,
                         ntiles, sigma, ocol, (uint16_t*)oval);
    else
      hipLaunchKernelGGL((k_c5_transpose<V, V, 0>), dim3(g), dim3(kWG), 0, s, col, val, nnz, ntiles,
                         sigma, ocol, (V*)oval);
    c5_launched("k_c5_transpose launch");
  }
  // 5. the kernels' image: row pointers, empty rows, segment ends, long list
  int32_t* orst = (int32_t*)M.put(nullptr, (size_t)(m + 1) * 4);
  if (int rc = hip_check(hipMemcpyAsync(orst, rst, (size_t)(m + 1) * 4, hipMemcpyDeviceToDevice, s),
                         "copy"))
    return rc;
  int32_t* elist = (int32_t*)M.put(nullptr, (size_t)total_empty * 4);
  hipLaunchKernelGGL(k_c5_empty_list, dim3(c5_grid(m)), dim3(kWG), 0, s, ef, nempty, m, elist);
  c5_launched("k_c5_empty_list launch");
  uint32_t tstart = 0;
  if (int rc = hip_check(hipMemcpyAsync(&tstart, tpf + p - 1, 4, hipMemcpyDeviceToHost, s), "read"))
    return rc;
  if (int rc = hip_check(hipStreamSynchronize(s), "read")) return rc;
  const int64_t tail_start = tstart & 0x7FFFFFFFu;
  int32_t rts = 0;
  if (tail_start < m) {
    if (int rc = hip_check(hipMemcpyAsync(&rts, rst + tail_start, 4, hipMemcpyDeviceToHost, s), "read"))
      return rc;
    if (int rc = hip_check(hipStreamSynchronize(s), "read")) return rc;
  }
  d.tail_start = (int)tail_start;
  d.first_tail_row = (tail_start < m && rts < d.tail_base) ? (int)tail_start + 1 : (int)tail_start;
  const int64_t ntail = m - d.first_tail_row;
  int32_t* chain = (int32_t*)M.put(nullptr, (size_t)ntiles * 4);
  int64_t* a = (int64_t*)M.temp((size_t)(5 * ntiles + 2 * std::max<int64_t>(ntail, 0)) * 8 + 8);
  int64_t *sm = a + ntiles, *lflag = sm + ntiles, *lpos = lflag + ntiles;
  int64_t *tflag = lpos + ntiles, *tpos = tflag + std::max<int64_t>(ntail, 0);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_c5_rev_stops, dim3(c5_grid(ntiles)), dim3(kWG), 0, s, kind, ntiles, a);
    c5_launched("k_c5_rev_stops launch");
    size_t tmp = 0;
    int rc = hip_check(rocprim::inclusive_scan(nullptr, tmp, a, sm, (size_t)ntiles,
                                               rocprim::minimum<int64_t>(), s),
                       "scan size");
    if (rc) return rc;
    void* tbuf = M.temp(tmp);
    rc = hip_check(rocprim::inclusive_scan(tbuf, tmp, a, sm, (size_t)ntiles,
                                           rocprim::minimum<int64_t>(), s),
                   "scan");
    if (rc) return rc;
    hipLaunchKernelGGL(k_c5_chain, dim3(c5_grid(ntiles)), dim3(kWG), 0, s, kind, sm, ntiles, chain,
                       lflag);
    c5_launched("k_c5_chain launch");
  }
  const int64_t nl = c5_scan(M, lflag, lpos, ntiles, s);
  int64_t nt = 0;
  if (ntail > 0) {
    hipLaunchKernelGGL(k_c5_tail_flags, dim3(c5_grid(ntail)), dim3(kWG), 0, s, rst,
                       (int64_t)d.first_tail_row, m, tflag);
    c5_launched("k_c5_tail_flags launch");
    nt = c5_scan(M, tflag, tpos, ntail, s);
  }
  int32_t* longs = (int32_t*)M.put(nullptr, (size_t)(nl + nt) * 4);
  if (nl + nt > 0) {
    hipLaunchKernelGGL(k_c5_long_list, dim3(c5_grid(ntiles + std::max<int64_t>(ntail, 0))),
                       dim3(kWG), 0, s, lflag, lpos, ntiles, tflag, tpos,
                       std::max<int64_t>(ntail, 0), nl, (int64_t)d.first_tail_row, longs);
    c5_launched("k_c5_long_list launch");
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device CSR5 build")) return rc;
  M.free_temps();
  const size_t acc = M.dtype == RVVLA_F64 ? 8 : 4;
  d.scratch = M.put(nullptr, (size_t)ntiles * (2 * acc + 4) + 16);
  M.device_bytes -= (int64_t)((size_t)ntiles * (2 * acc + 4) + 16);
  d.tile_ptr = tpf;
  d.desc = desc;
  d.eptr = eptr;
  d.eoff = eoff;
  d.noffsets = noff;
  d.rst = orst;
  d.col = ocol;
  d.val = oval;
  d.empty_rows = elist;
  d.nempty = total_empty;
  d.chain = chain;
  d.long_fix = longs;
  d.nlong = nl + nt;
  return RVVLA_OK;
}

template int build_csr5_device<double>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int, hipStream_t);
template int build_csr5_device<float>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int, hipStream_t);

}  // namespace rvvla
