// k_convert_lav.hip -- device-side CSR -> LAV (SIMD_Lanes = 64) builder.
//
// The layout of the host builder (convert_lav.cpp, after Gomez et al.,
// PPoPP 2021, in the fields the reference kernel reads, reference
// src/scalar/spmv_mtx_lav.cpp:52-79) from a CSR already on the device, with
// histogram, sort, scan and scatter passes:
//  1. column entry counts (atomic histogram); a stable radix sort by
//     descending count ranks the columns (ties by index); the dense part is
//     the shortest ranked prefix holding (1 - T) of the entries (binary
//     search over the scanned counts), cut into segments of segment_len
//     ranked columns;
//  2. sparse_part: the entries of the other columns, in CSR order (a scan
//     of per-entry flags gives every one its position and every row its
//     row pointer);
//  3. per (row, segment) dense counts (atomics); a stable radix sort of the
//     rows by descending dense length gives the row order; per (segment,
//     chunk) the longest and shortest of its 64 rows give the steps and the
//     full-mask steps, a scan of the steps the chunk offsets and segment
//     bases;
//  4. each row places its dense entries (lane = its position in the chunk,
//     step = running count within the segment) and pads its remaining slots
//     with its last dense column -- a wavefront per row in the chunks that
//     hold a row of more than 32 dense entries, a thread per row in the
//     others; one wavefront per (segment, chunk) writes the lane masks of
//     its partial steps from the suffix maxima of its lane counts.
// Then the kernels' image as the upload builds it (sparse_part as CSR with
// its row blocks, long chunks as pieces, the fused form when every
// sparse_part row has at most kLavFuseRow entries).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

constexpr int kLavL = 64;

__global__ void k_lv_colcount(const int32_t* __restrict__ col, int64_t nnz,
                              uint32_t* __restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e < nnz) atomicAdd(&cnt[col[e]], 1u);
}

__global__ void k_lv_colkeys(const uint32_t* __restrict__ cnt, int64_t n,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const int64_t c = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (c >= n) return;
  key[c] = 0xFFFFFFFFu - cnt[c];
  idx[c] = (uint32_t)c;
}

__global__ void k_lv_sorted_counts(const uint32_t* __restrict__ skey, int64_t n,
                                   int64_t* __restrict__ scnt) {
  const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (k < n) scnt[k] = (int64_t)(0xFFFFFFFFu - skey[k]);
}

// ndense: the first k with (double)(sum of the k largest counts) >= need,
// or the number of columns with entries if that is smaller
__global__ void k_lv_ndense(const int64_t* __restrict__ pre, const int64_t* __restrict__ scnt,
                            int64_t n, double need, int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t lo = 0, hi = n;  // smallest k in [0, n] with P(k) >= need, P(0) = 0, P(k) = pre[k-1]
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int64_t P = mid == 0 ? 0 : pre[mid - 1];
    if ((double)P >= need) hi = mid;
    else lo = mid + 1;
  }
  const int64_t k1 = lo;
  lo = 0;
  hi = n;  // first k with scnt[k] == 0 (counts descend)
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (scnt[mid] == 0) hi = mid;
    else lo = mid + 1;
  }
  out[0] = k1 < lo ? k1 : lo;
}

__global__ void k_lv_segof(const uint32_t* __restrict__ order, int64_t ndense, int64_t seg_len,
                           int32_t* __restrict__ seg_of) {
  const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (k < ndense) seg_of[order[k]] = (int32_t)(k / seg_len);
}

__global__ void k_lv_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_lv_fill_u16(uint16_t* __restrict__ p, int64_t n, uint16_t v) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < n) p[i] = v;
}

// per entry: sparse flag; per dense entry: its (row, segment) count and the
// row's dense length
__global__ void k_lv_entry_counts(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                                  const int32_t* __restrict__ seg_of, int64_t m, int64_t nnz,
                                  int64_t S, int64_t* __restrict__ spf,
                                  int32_t* __restrict__ rcnt, int32_t* __restrict__ dlen) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e >= nnz) return;
  const int32_t g = seg_of[col[e]];
  spf[e] = g < 0 ? 1 : 0;
  if (g < 0) return;
  int64_t lo = 0, hi = m;  // the row of e: the last r with rst[r] <= e
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rst[mid] <= e) lo = mid;
    else hi = mid;
  }
  atomicAdd(&rcnt[lo * S + g], 1);
  atomicAdd(&dlen[lo], 1);
}

template <class V, class SV, int NARROW>
__device__ __forceinline__ void lv_store(SV* p, int64_t i, V v) {
  if constexpr (NARROW == 1) p[i] = dev_bf16((float)v);
  else if constexpr (NARROW == 2) p[i] = dev_f16((float)v);
  else p[i] = v;
}

template <class V, class SV, int NARROW>
__global__ void k_lv_sparse(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                            const V* __restrict__ val, const int64_t* __restrict__ spf,
                            const int64_t* __restrict__ spos, int64_t m, int64_t nnz, int64_t nsp,
                            int32_t* __restrict__ srst, int32_t* __restrict__ scol,
                            SV* __restrict__ sval) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < nnz && spf[i]) {
    scol[spos[i]] = col[i];
    lv_store<V, SV, NARROW>(sval, spos[i], val[i]);
  }
  if (i <= m) srst[i] = rst[i] < nnz ? (int32_t)spos[rst[i]] : (int32_t)nsp;
}

__global__ void k_lv_rowkeys(const int32_t* __restrict__ dlen, int64_t m,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r >= m) return;
  key[r] = 0xFFFFFFFFu - (uint32_t)dlen[r];
  idx[r] = (uint32_t)r;
}

__global__ void k_lv_out_order(const uint32_t* __restrict__ perm, int64_t m, int64_t span,
                               int64_t total, uint32_t* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (j >= total) return;
  const int64_t i = j % span;
  out[j] = i < m ? perm[i] : perm[m - 1];
}

// steps and full-mask steps of chunk c in segment g (q = g * cc + c)
__global__ void k_lv_steps(const int32_t* __restrict__ rcnt, const uint32_t* __restrict__ perm,
                           int64_t m, int64_t S, int64_t cc, int64_t* __restrict__ steps,
                           int64_t* __restrict__ full) {
  const int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= S * cc) return;
  const int64_t g = q / cc, c = q % cc;
  int64_t mx = 0, mn = INT64_MAX;
  for (int i = 0; i < kLavL; ++i) {
    const int64_t rp = c * kLavL + i;
    const int64_t k = rp < m ? rcnt[(int64_t)perm[rp] * S + g] : 0;
    mx = k > mx ? k : mx;
    mn = k < mn ? k : mn;
  }
  steps[q] = mx;
  full[q] = mn;
}

__global__ void k_lv_offsets(const int64_t* __restrict__ G, const int64_t* __restrict__ full,
                             int64_t S, int64_t cc, int64_t total,
                             int32_t* __restrict__ coff, int32_t* __restrict__ cfull,
                             int64_t* __restrict__ sptr) {
  const int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= S * (cc + 1)) return;
  const int64_t g = q / (cc + 1), c = q % (cc + 1);
  const int64_t base = G[g * cc];
  const int64_t end = g + 1 < S ? G[(g + 1) * cc] : total;
  const int64_t off = c < cc ? G[g * cc + c] - base : end - base;
  coff[q] = (int32_t)off;
  if (c < cc) cfull[g * cc + c] = (int32_t)(off + full[g * cc + c]);
  if (c == 0) sptr[g] = base;
  if (g == S - 1 && c == cc) sptr[S] = total;
}

// one thread per row (sorted rows rp0 .. m-1, in chunks whose rows have at
// most kLvShortRow dense entries): its dense entries into its lane, then
// the pads
constexpr int kLvShortRow = 32;
template <class V, class SV, int NARROW>
__global__ void k_lv_place(const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
                           const V* __restrict__ val, const int32_t* __restrict__ seg_of,
                           const uint32_t* __restrict__ perm, const int32_t* __restrict__ coff,
                           const int64_t* __restrict__ sptr, int64_t m, int64_t S, int64_t cc,
                           int64_t rp0, uint32_t* __restrict__ col_id, SV* __restrict__ dval) {
  const int64_t rp = rp0 + (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (rp >= m) return;
  const uint32_t r = perm[rp];
  const int64_t c = rp / kLavL, i = rp % kLavL;
  const int e0 = rst[r], e1 = rst[r + 1];
  uint32_t last_col = 0;
  for (int e = e0; e < e1; ++e)
    if (seg_of[col[e]] >= 0) last_col = (uint32_t)col[e];
  for (int64_t g = 0; g < S; ++g) {
    int64_t pos = sptr[g] + coff[g * (cc + 1) + c];
    const int64_t end = sptr[g] + coff[g * (cc + 1) + c + 1];
    for (int e = e0; e < e1; ++e) {
      if (seg_of[col[e]] != g) continue;
      col_id[pos * kLavL + i] = (uint32_t)col[e];
      lv_store<V, SV, NARROW>(dval, pos * kLavL + i, val[e]);
      ++pos;
    }
    for (; pos < end; ++pos) col_id[pos * kLavL + i] = last_col;
  }
}

// one wavefront per sorted row rp < rp_end (the chunks holding rows with
// more than kLvShortRow dense entries): per segment the row's entries of
// that segment in entry order (positions by ballot and prefix count), then
// the pads with the row's last dense column
template <class V, class SV, int NARROW>
__global__ __launch_bounds__(kWG) void k_lv_place_wave(
    const int32_t* __restrict__ rst, const int32_t* __restrict__ col, const V* __restrict__ val,
    const int32_t* __restrict__ seg_of, const uint32_t* __restrict__ perm,
    const int32_t* __restrict__ coff, const int64_t* __restrict__ sptr, int64_t rp_end, int64_t S,
    int64_t cc, uint32_t* __restrict__ col_id, SV* __restrict__ dval) {
  const int64_t rp = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (rp >= rp_end) return;
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t r = perm[rp];
  const int64_t c = rp / kLavL, i = rp % kLavL;
  const int e0 = rst[r], e1 = rst[r + 1];
  int best = -1;  // the row's last dense entry
  for (int e = e0 + lane; e < e1; e += 64)
    if (seg_of[col[e]] >= 0) best = e;
  for (int d = 32; d > 0; d >>= 1) {
    const int o = __shfl_xor(best, d, 64);
    best = o > best ? o : best;
  }
  const uint32_t last_col = best >= 0 ? (uint32_t)col[best] : 0u;
  for (int64_t g = 0; g < S; ++g) {
    const int64_t pos0 = sptr[g] + coff[g * (cc + 1) + c];
    const int64_t end = sptr[g] + coff[g * (cc + 1) + c + 1];
    int64_t run = 0;
    for (int b = e0; b < e1; b += 64) {
      const int e = b + lane;
      const int32_t cl = e < e1 ? col[e] : 0;
      const bool mine = e < e1 && seg_of[cl] == g;
      const uint64_t bal = __ballot(mine);
      if (mine) {
        const int64_t p = pos0 + run + __popcll(bal & below);
        col_id[p * kLavL + i] = (uint32_t)cl;
        lv_store<V, SV, NARROW>(dval, p * kLavL + i, val[e]);
      }
      run += __popcll(bal);
    }
    for (int64_t p = pos0 + run + lane; p < end; p += 64) col_id[p * kLavL + i] = last_col;
  }
}

// lane masks of the partial steps of chunk c in segment g, one wavefront
// per (g, c): with M_j the largest count of lanes j .. 63, the mask of
// step k (t = k - offset) is the last lane with a count above t, plus one,
// = the number of j with M_j > t (M is non-increasing in j)
__global__ __launch_bounds__(kWG) void k_lv_masks(
    const int32_t* __restrict__ rcnt, const uint32_t* __restrict__ perm,
    const int32_t* __restrict__ coff, const int32_t* __restrict__ cfull,
    const int64_t* __restrict__ sptr, int64_t m, int64_t S, int64_t cc,
    uint16_t* __restrict__ mask) {
  __shared__ int32_t sm[kWG];
  const int64_t q = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  int32_t* M = sm + (threadIdx.x & ~63);
  const bool live = q < S * cc;
  const int64_t g = live ? q / cc : 0, c = live ? q % cc : 0;
  const int64_t rp = c * kLavL + lane;
  int32_t v = live && rp < m ? rcnt[(int64_t)perm[rp] * S + g] : 0;
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t o = __shfl_down(v, d, 64);
    if (lane + d < 64) v = o > v ? o : v;
  }
  M[lane] = v;
  __syncthreads();
  if (!live) return;
  const int64_t off = coff[g * (cc + 1) + c], b = coff[g * (cc + 1) + c + 1];
  for (int64_t k = cfull[q] + lane; k < b; k += 64) {
    const int64_t t = k - off;
    int lo = 0, hi = 64;  // the first j with M[j] <= t
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)M[mid] > t) lo = mid + 1;
      else hi = mid;
    }
    mask[sptr[g] + k] = (uint16_t)lo;
  }
}

// the number of rows with more than kLvShortRow dense entries (sorted keys
// are 0xFFFFFFFF - dense length, ascending)
__global__ void k_lv_count_long(const uint32_t* __restrict__ skey, int64_t m,
                                int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t lim = 0xFFFFFFFFu - (uint32_t)kLvShortRow;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (skey[mid] < lim) lo = mid + 1;
    else hi = mid;
  }
  *out = lo;
}

namespace {

unsigned lv_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

void lv_launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

template <class OP>
void lv_scan(rvvla_mat_s& M, const int64_t* in, int64_t* out, int64_t n, bool inclusive, OP op,
             hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  int rc = inclusive
               ? hip_check(rocprim::inclusive_scan(nullptr, tmp, in, out, (size_t)n, op, s), "scan size")
               : hip_check(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0, (size_t)n, op, s),
                           "scan size");
  if (rc) throw Error(rc, last_error());
  void* tbuf = M.temp(tmp);
  rc = inclusive
           ? hip_check(rocprim::inclusive_scan(tbuf, tmp, in, out, (size_t)n, op, s), "scan")
           : hip_check(rocprim::exclusive_scan(tbuf, tmp, in, out, (int64_t)0, (size_t)n, op, s),
                       "scan");
  if (rc) throw Error(rc, last_error());
}

int64_t lv_read64(const int64_t* p, hipStream_t s) {
  int64_t v = 0;
  int rc = hip_check(hipMemcpyAsync(&v, p, 8, hipMemcpyDeviceToHost, s), "read");
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "read");
  if (rc) throw Error(rc, last_error());
  return v;
}

void lv_sort_pairs(rvvla_mat_s& M, uint32_t* key, uint32_t* key_out, uint32_t* val,
                   uint32_t* val_out, int64_t n, hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  int rc = hip_check(rocprim::radix_sort_pairs(nullptr, tmp, key, key_out, val, val_out,
                                               (size_t)n, 0, 32, s),
                     "radix sort size");
  if (rc) throw Error(rc, last_error());
  void* tbuf = M.temp(tmp);
  rc = hip_check(rocprim::radix_sort_pairs(tbuf, tmp, key, key_out, val, val_out, (size_t)n, 0,
                                           32, s),
                 "radix sort");
  if (rc) throw Error(rc, last_error());
}

}  // namespace

template <class V>
int build_lav_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                     const int32_t* col, const V* val, double T, int64_t seg_len,
                     LavDeviceHost& host, hipStream_t s) {
  const int64_t cc = (m + kLavL - 1) / kLavL;
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  const int narrow = M.dtype == RVVLA_BF16 ? 1 : M.dtype == RVVLA_F16 ? 2 : 0;
  // 1. column ranking, dense part, segments
  uint32_t* cnt = (uint32_t*)M.temp((size_t)n * 20 + 8);
  uint32_t *ckey = cnt + n, *cidx = ckey + n, *skey = cidx + n, *order = skey + n;
  if (int rc = hip_check(hipMemsetAsync(cnt, 0, (size_t)n * 4, s), "clear")) return rc;
  if (nnz > 0) {
    hipLaunchKernelGGL(k_lv_colcount, dim3(lv_grid(nnz)), dim3(kWG), 0, s, col, nnz, cnt);
    lv_launched("k_lv_colcount launch");
  }
  hipLaunchKernelGGL(k_lv_colkeys, dim3(lv_grid(n)), dim3(kWG), 0, s, cnt, n, ckey, cidx);
  lv_launched("k_lv_colkeys launch");
  lv_sort_pairs(M, ckey, skey, cidx, order, n, s);
  int64_t* scnt = (int64_t*)M.temp((size_t)n * 16 + 16);
  int64_t *pre = scnt + n, *nd = pre + n;
  hipLaunchKernelGGL(k_lv_sorted_counts, dim3(lv_grid(n)), dim3(kWG), 0, s, skey, n, scnt);
  lv_launched("k_lv_sorted_counts launch");
  lv_scan(M, scnt, pre, n, true, rocprim::plus<int64_t>(), s);
  const double need = (1.0 - T) * (double)nnz;
  hipLaunchKernelGGL(k_lv_ndense, dim3(1), dim3(64), 0, s, pre, scnt, n, need, nd);
  lv_launched("k_lv_ndense launch");
  const int64_t ndense = n > 0 ? lv_read64(nd, s) : 0;
  const int64_t S = (ndense + seg_len - 1) / seg_len;
  int32_t* seg_of = (int32_t*)M.temp((size_t)n * 4 + 4);
  hipLaunchKernelGGL(k_lv_fill_i32, dim3(lv_grid(n)), dim3(kWG), 0, s, seg_of, n, -1);
  lv_launched("k_lv_fill_i32 launch");
  if (ndense > 0) {
    hipLaunchKernelGGL(k_lv_segof, dim3(lv_grid(ndense)), dim3(kWG), 0, s, order, ndense, seg_len,
                       seg_of);
    lv_launched("k_lv_segof launch");
  }
  // 2-3. sparse_part, per (row, segment) counts, row order
  int64_t* spf = (int64_t*)M.temp((size_t)nnz * 16 + 8);
  int64_t* spos = spf + nnz;
  int32_t* rcnt = (int32_t*)M.temp((size_t)(m * std::max<int64_t>(S, 1) + m) * 4 + 4);
  int32_t* dlen = rcnt + m * std::max<int64_t>(S, 1);
  if (int rc = hip_check(hipMemsetAsync(rcnt, 0, (size_t)(m * std::max<int64_t>(S, 1) + m) * 4, s),
                         "clear"))
    return rc;
  if (nnz > 0) {
    hipLaunchKernelGGL(k_lv_entry_counts, dim3(lv_grid(nnz)), dim3(kWG), 0, s, rst, col, seg_of, m,
                       nnz, std::max<int64_t>(S, 1), spf, rcnt, dlen);
    lv_launched("k_lv_entry_counts launch");
  }
  lv_scan(M, spf, spos, nnz, false, rocprim::plus<int64_t>(), s);
  const int64_t nsp = nnz > 0 ? lv_read64(spos + nnz - 1, s) + lv_read64(spf + nnz - 1, s) : 0;
  int32_t* srst = (int32_t*)M.put(nullptr, (size_t)(m + 1) * 4);
  int32_t* scol = (int32_t*)M.put(nullptr, (size_t)nsp * 4);
  void* sval = M.put(nullptr, (size_t)nsp * vbytes);
  {
    const unsigned g = lv_grid(std::max<int64_t>(nnz, m + 1));
    if (narrow == 1)
      hipLaunchKernelGGL((k_lv_sparse<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, rst, col, val,
                         spf, spos, m, nnz, nsp, srst, scol, (uint16_t*)sval);
    else if (narrow == 2)
      hipLaunchKernelGGL((k_lv_sparse<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, rst, col, val,
                         spf, spos, m, nnz, nsp, srst, scol, (uint16_t*)sval);
    else
      hipLaunchKernelGGL((k_lv_sparse<V, V, 0>), dim3(g), dim3(kWG), 0, s, rst, col, val, spf,
                         spos, m, nnz, nsp, srst, scol, (V*)sval);
    lv_launched("k_lv_sparse launch");
  }
  host.S = S;
  host.cc = cc;
  host.srst = srst;
  host.scol = scol;
  host.sval = sval;
  host.nsp = nsp;
  if (S == 0 || m == 0) {  // no dense part: out_order and the chunk tables stay empty
    host.total = 0;
    if (int rc = hip_check(hipStreamSynchronize(s), "device LAV build")) return rc;
    return RVVLA_OK;
  }
  uint32_t* rkey = (uint32_t*)M.temp((size_t)m * 16 + 8);
  uint32_t *ridx = rkey + m, *rkey2 = ridx + m, *perm = rkey2 + m;
  hipLaunchKernelGGL(k_lv_rowkeys, dim3(lv_grid(m)), dim3(kWG), 0, s, dlen, m, rkey, ridx);
  lv_launched("k_lv_rowkeys launch");
  lv_sort_pairs(M, rkey, rkey2, ridx, perm, m, s);
  uint32_t* out_order = (uint32_t*)M.put(nullptr, (size_t)(S * cc * kLavL) * 4);
  hipLaunchKernelGGL(k_lv_out_order, dim3(lv_grid(S * cc * kLavL)), dim3(kWG), 0, s, perm, m,
                     cc * kLavL, S * cc * kLavL, out_order);
  lv_launched("k_lv_out_order launch");
  int64_t* steps = (int64_t*)M.temp((size_t)(S * cc) * 24 + 8);
  int64_t *full = steps + S * cc, *G = full + S * cc;
  hipLaunchKernelGGL(k_lv_steps, dim3(lv_grid(S * cc)), dim3(kWG), 0, s, rcnt, perm, m, S, cc,
                     steps, full);
  lv_launched("k_lv_steps launch");
  lv_scan(M, steps, G, S * cc, false, rocprim::plus<int64_t>(), s);
  const int64_t total = lv_read64(G + S * cc - 1, s) + lv_read64(steps + S * cc - 1, s);
  if (total * kLavL > 2147483647LL) {
    set_last_error("LAV: dense part too large");
    return RVVLA_ERR_LENGTH;
  }
  int32_t* coff = (int32_t*)M.put(nullptr, (size_t)(S * (cc + 1)) * 4);
  int32_t* cfull = (int32_t*)M.put(nullptr, (size_t)(S * cc) * 4);
  int64_t* sptr = (int64_t*)M.put(nullptr, (size_t)(S + 1) * 8);
  hipLaunchKernelGGL(k_lv_offsets, dim3(lv_grid(S * (cc + 1))), dim3(kWG), 0, s, G, full, S, cc,
                     total, coff, cfull, sptr);
  lv_launched("k_lv_offsets launch");
  // 4. entries, pads, masks
  uint32_t* col_id = (uint32_t*)M.put(nullptr, (size_t)(total * kLavL) * 4);
  void* dval = M.put(nullptr, (size_t)(total * kLavL) * vbytes);
  uint16_t* mask = (uint16_t*)M.put(nullptr, (size_t)total * 2);
  if (int rc = hip_check(hipMemsetAsync(col_id, 0, (size_t)(total * kLavL) * 4, s), "clear")) return rc;
  if (int rc = hip_check(hipMemsetAsync(dval, 0, (size_t)(total * kLavL) * vbytes, s), "clear")) return rc;
  if (total > 0) {
    hipLaunchKernelGGL(k_lv_fill_u16, dim3(lv_grid(total)), dim3(kWG), 0, s, mask, total,
                       (uint16_t)kLavL);
    lv_launched("k_lv_fill_u16 launch");
  }
  // rows of the chunks that hold a row of more than kLvShortRow dense
  // entries: a wavefront each; the rest: a thread each
  int64_t* nlong = steps;  // reused: its scan is done
  hipLaunchKernelGGL(k_lv_count_long, dim3(1), dim3(64), 0, s, rkey2, m, nlong);
  lv_launched("k_lv_count_long launch");
  const int64_t rp0 = std::min<int64_t>(m, (lv_read64(nlong, s) + kLavL - 1) / kLavL * kLavL);
  if (rp0 > 0) {
    const unsigned g = (unsigned)((rp0 + kWG / 64 - 1) / (kWG / 64));
    if (narrow == 1)
      hipLaunchKernelGGL((k_lv_place_wave<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, rst, col,
                         val, seg_of, perm, coff, sptr, rp0, S, cc, col_id, (uint16_t*)dval);
    else if (narrow == 2)
      hipLaunchKernelGGL((k_lv_place_wave<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, rst, col,
                         val, seg_of, perm, coff, sptr, rp0, S, cc, col_id, (uint16_t*)dval);
    else
      hipLaunchKernelGGL((k_lv_place_wave<V, V, 0>), dim3(g), dim3(kWG), 0, s, rst, col, val,
                         seg_of, perm, coff, sptr, rp0, S, cc, col_id, (V*)dval);
    lv_launched("k_lv_place_wave launch");
  }
  if (rp0 < m) {
    const unsigned g = lv_grid(m - rp0);
    if (narrow == 1)
      hipLaunchKernelGGL((k_lv_place<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, rst, col, val,
                         seg_of, perm, coff, sptr, m, S, cc, rp0, col_id, (uint16_t*)dval);
    else if (narrow == 2)
      hipLaunchKernelGGL((k_lv_place<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, rst, col, val,
                         seg_of, perm, coff, sptr, m, S, cc, rp0, col_id, (uint16_t*)dval);
    else
      hipLaunchKernelGGL((k_lv_place<V, V, 0>), dim3(g), dim3(kWG), 0, s, rst, col, val, seg_of,
                         perm, coff, sptr, m, S, cc, rp0, col_id, (V*)dval);
    lv_launched("k_lv_place launch");
  }
  hipLaunchKernelGGL(k_lv_masks, dim3((unsigned)((S * cc + kWG / 64 - 1) / (kWG / 64))), dim3(kWG),
                     0, s, rcnt, perm, coff, cfull, sptr, m, S, cc, mask);
  lv_launched("k_lv_masks launch");
  if (int rc = hip_check(hipStreamSynchronize(s), "device LAV build")) return rc;
  host.total = total;
  host.out_order = out_order;
  host.coff = coff;
  host.cfull = cfull;
  host.sptr = sptr;
  host.col_id = col_id;
  host.dval = dval;
  host.mask = mask;
  return RVVLA_OK;
}

template int build_lav_device<double>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const double*, double, int64_t,
                                      LavDeviceHost&, hipStream_t);
template int build_lav_device<float>(rvvla_mat_s&, int64_t, int64_t, int64_t, const int32_t*,
                                     const int32_t*, const float*, double, int64_t,
                                     LavDeviceHost&, hipStream_t);

}  // namespace rvvla
