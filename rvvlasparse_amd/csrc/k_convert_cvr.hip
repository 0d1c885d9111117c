// k_convert_cvr.hip -- device-side CSR -> CVR (SIMD_LEN = 64) builder.
//
// The device image that convert(RVVLA_FMT_CVR) + upload give (the layout
// of convert_cvr.cpp after Xie et al., CGO 2018), built from a CSR already
// on the device:
//  1. the segments: every non-empty row cut into pieces of at most
//     cvr_segment_cap entries (a count per row, a scan, one thread per row
//     writing its segments' starts and row), and the list of empty rows;
//  2. the balanced stream boundaries over the segments (lower bound of each
//     stream's entry target in the segment starts, one thread each) and,
//     per stream, the lower bound on its packed slots (refused as on the
//     host when out of range);
//  3. one wavefront per stream replays the greedy packing: a lane whose
//     segment is used up records its row and takes the next segment -- the
//     finishing lanes take segments in lane order, so lane l's new one is
//     the one at the popcount of the finishing lanes below it.  A first
//     pass counts packs and final-flush records; scans place every stream;
//     a second pass writes the packs, the flush masks, the records and the
//     final-flush rows;
//  4. rows recorded more than once (cut rows): every flush of such a row,
//     in record order, gets a partial-sum slot, the rows' slot ranges
//     following one another in row order -- a stable radix sort of the
//     flushes by row gives each its slot as its sorted position -- and the
//     rows are listed for the combine (more than 64 slots first).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

constexpr int kVcWaves = kWG / 64;

__global__ void k_vc_parts(const int32_t* __restrict__ rst, int64_t m, int64_t nnz, int64_t S,
                           int32_t* __restrict__ rb) {  // rst: the segment starts (m segments)
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > S) return;
  if (t == 0 || t == S) {
    rb[t] = t == 0 ? 0 : (int32_t)m;
    return;
  }
  int64_t target = (nnz * t + S - 1) / S;
  if (target > nnz) target = nnz;
  int64_t lo = 0, hi = m + 1;  // the first i with rst[i] >= target
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)rst[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  rb[t] = (int32_t)lo;
}

__global__ void k_vc_flags(const int32_t* __restrict__ rst, int64_t m, int64_t* __restrict__ f) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r <= m) f[r] = r < m && rst[r + 1] > rst[r] ? 1 : 0;
}

// the empty rows in order (nep: the non-empty rows before each row)
__global__ void k_vc_empty(const int32_t* __restrict__ rst, const int64_t* __restrict__ nep,
                           int64_t m, int32_t* __restrict__ empty) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m && rst[r + 1] == rst[r]) empty[r - nep[r]] = (int32_t)r;
}

__device__ __forceinline__ int64_t vc_nseg(const int32_t* rst, int64_t r, int64_t cap) {
  const int64_t len = (int64_t)rst[r + 1] - rst[r];
  return (len + cap - 1) / cap;
}

__global__ void k_vc_nseg(const int32_t* __restrict__ rst, int64_t m, int64_t cap,
                          int64_t* __restrict__ nseg) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m) nseg[r] = vc_nseg(rst, r, cap);
}

// segment starts (vrst, nv + 1) and rows (vrow)
__global__ void k_vc_segs(const int32_t* __restrict__ rst, int64_t m, int64_t cap, int64_t nv,
                          const int64_t* __restrict__ vbase, int32_t* __restrict__ vrst,
                          int32_t* __restrict__ vrow) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r > m) return;
  if (r == m) {
    vrst[nv] = m > 0 ? rst[m] : 0;
    return;
  }
  const int64_t n = vc_nseg(rst, r, cap);
  for (int64_t k = 0; k < n; ++k) {
    vrst[vbase[r] + k] = (int32_t)(rst[r] + k * cap);
    vrow[vbase[r] + k] = (int32_t)r;
  }
}

// one wavefront per stream: max(ceil(entries / 64), longest segment) * 64
__global__ __launch_bounds__(kWG) void k_vc_bound(const int32_t* __restrict__ rst,
                                                  const int32_t* __restrict__ rb, int64_t S,
                                                  int64_t* __restrict__ bound) {
  const int64_t t = (int64_t)blockIdx.x * kVcWaves + (threadIdx.x >> 6);
  if (t >= S) return;
  const int lane = threadIdx.x & 63;
  const int64_t r0 = rb[t], r1 = rb[t + 1];
  int64_t longest = 0;
  for (int64_t r = r0 + lane; r < r1; r += 64) {
    const int64_t len = (int64_t)rst[r + 1] - rst[r];
    longest = len > longest ? len : longest;
  }
  for (int d = 32; d > 0; d >>= 1) {
    const int64_t o = __shfl_xor(longest, d, 64);
    longest = o > longest ? o : longest;
  }
  if (lane == 0) {
    const int64_t nz = (int64_t)rst[r1] - rst[r0];
    const int64_t packs = (nz + 63) / 64;
    bound[t] = (packs > longest ? packs : longest) * 64;
  }
}

template <class V, class S, int NARROW, bool WRITE>
__global__ __launch_bounds__(kWG) void k_vc_pack(
    const int32_t* __restrict__ vrst, const int32_t* __restrict__ vrow,
    const int32_t* __restrict__ col, const V* __restrict__ val, const int32_t* __restrict__ rb,
    int64_t nstreams, int64_t* __restrict__ npk, int64_t* __restrict__ nrec_out,
    const int64_t* __restrict__ pstart, const int64_t* __restrict__ rstart,
    uint64_t* __restrict__ flush, int32_t* __restrict__ rec_rows, int32_t* __restrict__ tail_rows,
    uint32_t* __restrict__ ocol, S* __restrict__ oval) {
  const int64_t t = (int64_t)blockIdx.x * kVcWaves + (threadIdx.x >> 6);
  if (t >= nstreams) return;
  const int lane = threadIdx.x & 63;
  const int64_t nend = rb[t + 1];
  int64_t next = rb[t];
  uint32_t pad = 0;  // the column of the stream's last entry
  if (WRITE && nend > next) pad = (uint32_t)col[vrst[nend] - 1];
  int32_t row = -1;
  int64_t pos = 0, stop = 0;
  auto take = [&](int64_t idx) {  // segment idx, if the stream has it
    if (idx < nend) {
      row = vrow[idx];
      pos = vrst[idx];
      stop = vrst[idx + 1];
    } else {
      row = -1;
    }
  };
  take(next + lane);
  next = next + 64 < nend ? next + 64 : nend;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int64_t nrec = 0;
  for (int64_t k = 0;; ++k) {
    const bool f = row >= 0 && pos == stop;
    const uint64_t fin = __ballot(f);
    const int rank = __popcll(fin & below);
    const int nfin = __popcll(fin);
    const int32_t done_row = row;
    if (f) take(next + rank);
    next = next + nfin < nend ? next + nfin : nend;
    if (__ballot(row >= 0) == 0) {  // every lane idle: the final flush
      if (WRITE) tail_rows[t * 64 + lane] = f ? done_row : -1;
      else if (lane == 0) {
        npk[t] = k;
        nrec_out[t] = nrec;
      }
      break;
    }
    if (WRITE) {
      const int64_t kk = pstart[t] + k;
      if (lane == 0) flush[kk] = fin;
      if (f) rec_rows[rstart[t] + nrec + rank] = done_row;
      const int64_t slot = kk * 64 + lane;
      V v = V(0);
      uint32_t c = pad;
      if (row >= 0) {
        v = val[pos];
        c = (uint32_t)col[pos];
      }
      ocol[slot] = c;
      if constexpr (NARROW == 1) oval[slot] = dev_bf16((float)v);
      else if constexpr (NARROW == 2) oval[slot] = dev_f16((float)v);
      else oval[slot] = v;
    }
    nrec += nfin;
    if (row >= 0) ++pos;
  }
}

// flush g in record order (stream t: its records, then its 64 final
// flushes in lane order; g = record index + 64 t, or end of t's records +
// 64 t + lane): key = the row when it is recorded more than once (length
// above the cap), else ~0; loc = the record index, or -(final flush + 1)
__global__ void k_vc_flush_keys(const int32_t* __restrict__ rst, const int64_t* __restrict__ rs,
                                int64_t S, int64_t nrec, int64_t cap,
                                const int32_t* __restrict__ rec_rows,
                                const int32_t* __restrict__ tails, uint32_t* __restrict__ key,
                                uint32_t* __restrict__ gidx, int64_t* __restrict__ loc) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i >= nrec + 64 * S) return;
  int64_t g, where;
  int32_t row;
  if (i < nrec) {  // record i: its stream is the last t with rs[t] <= i
    int64_t lo = 0, hi = S;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (rs[mid] <= i) lo = mid;
      else hi = mid;
    }
    g = i + 64 * lo;
    where = i;
    row = rec_rows[i];
  } else {
    const int64_t f = i - nrec, t = f / 64, lane = f % 64;
    const int64_t end = t + 1 < S ? rs[t + 1] : nrec;
    g = end + 64 * t + lane;
    where = -(f + 1);
    row = tails[f];
  }
  const bool cut = row >= 0 && (int64_t)rst[row + 1] - rst[row] > cap;
  key[g] = cut ? (uint32_t)row : ~0u;
  gidx[g] = (uint32_t)g;
  loc[g] = where;
}

// sorted flush i of a cut row: its slot is i
__global__ void k_vc_flush_slots(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sg,
                                 const int64_t* __restrict__ loc, int64_t n,
                                 int32_t* __restrict__ rec_rows, int32_t* __restrict__ tails) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i >= n || skey[i] == ~0u) return;
  const int64_t w = loc[sg[i]];
  if (w >= 0) rec_rows[w] = cvr_slot_target(i);
  else tails[-w - 1] = cvr_slot_target(i);
}

// cut rows: slot counts (row order) and the long / short flags
__global__ void k_vc_cut_flags(const int32_t* __restrict__ rst, int64_t m, int64_t cap,
                               int64_t* __restrict__ slots, int64_t* __restrict__ longf,
                               int64_t* __restrict__ shortf) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r >= m) return;
  const int64_t n = vc_nseg(rst, r, cap);
  slots[r] = n > 1 ? n : 0;
  longf[r] = n > 64 ? 1 : 0;
  shortf[r] = n > 1 && n <= 64 ? 1 : 0;
}

__global__ void k_vc_cut_list(const int32_t* __restrict__ rst, int64_t m, int64_t cap,
                              const int64_t* __restrict__ sbase, const int64_t* __restrict__ li,
                              const int64_t* __restrict__ si, int64_t nlong,
                              int32_t* __restrict__ rows, int32_t* __restrict__ ptr) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r >= m) return;
  const int64_t n = vc_nseg(rst, r, cap);
  if (n <= 1) return;
  const int64_t at = n > 64 ? li[r] : nlong + si[r];
  rows[at] = (int32_t)r;
  ptr[2 * at] = (int32_t)sbase[r];
  ptr[2 * at + 1] = (int32_t)(sbase[r] + n);
}

__global__ void k_vc_starts(const int64_t* __restrict__ ps, const int64_t* __restrict__ rs,
                            int64_t S, int64_t npacks, int64_t* __restrict__ pack_start,
                            int32_t* __restrict__ rec_start) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > S) return;
  pack_start[t] = t < S ? ps[t] : npacks;
  if (t < S) rec_start[t] = (int32_t)rs[t];
}

namespace {

unsigned vc_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

unsigned vc_wave_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, (n + kVcWaves - 1) / kVcWaves);
}

void vc_launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

// exclusive scan of n counts into out; returns their total
int64_t vc_scan(rvvla_mat_s& M, const int64_t* cnt, int64_t* out, int64_t n, hipStream_t s) {
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

template <class V, class S, int NARROW>
void vc_write(const int32_t* vrst, const int32_t* vrow, const int32_t* col, const V* val,
              const int32_t* rb, int64_t nstreams, const int64_t* ps, const int64_t* rs,
              uint64_t* flush, int32_t* rec_rows, int32_t* tails, uint32_t* ocol, void* oval,
              hipStream_t s) {
  hipLaunchKernelGGL((k_vc_pack<V, S, NARROW, true>), dim3(vc_wave_grid(nstreams)), dim3(kWG), 0,
                     s, vrst, vrow, col, val, rb, nstreams, nullptr, nullptr, ps, rs, flush,
                     rec_rows, tails, ocol, (S*)oval);
  vc_launched("k_vc_pack launch");
}

}  // namespace

template <class V>
int build_cvr_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                     const int32_t* col, const V* val, int64_t nstreams, hipStream_t s) {
  DevCVR& d = M.cvr;
  const int64_t S = nstreams;
  const int64_t cap = cvr_segment_cap(nnz, S);
  // 1. segments and empty rows
  int64_t* nseg = (int64_t*)M.temp((size_t)(m + 1) * 16);
  int64_t* vbase = nseg + (m + 1);
  hipLaunchKernelGGL(k_vc_nseg, dim3(vc_grid(m)), dim3(kWG), 0, s, rst, m, cap, nseg);
  vc_launched("k_vc_nseg launch");
  const int64_t nv = vc_scan(M, nseg, vbase, m, s);
  int32_t* vrst = (int32_t*)M.temp((size_t)(nv + 1) * 4);
  int32_t* vrow = (int32_t*)M.temp((size_t)nv * 4 + 4);
  hipLaunchKernelGGL(k_vc_segs, dim3(vc_grid(m + 1)), dim3(kWG), 0, s, rst, m, cap, nv, vbase,
                     vrst, vrow);
  vc_launched("k_vc_segs launch");
  int64_t* f = (int64_t*)M.temp((size_t)(m + 1) * 16);
  int64_t* nep = f + (m + 1);
  hipLaunchKernelGGL(k_vc_flags, dim3(vc_grid(m + 1)), dim3(kWG), 0, s, rst, m, f);
  vc_launched("k_vc_flags launch");
  const int64_t nne = vc_scan(M, f, nep, m + 1, s);
  int32_t* elist = (int32_t*)M.put(nullptr, (size_t)(m - nne) * 4);
  hipLaunchKernelGGL(k_vc_empty, dim3(vc_grid(m)), dim3(kWG), 0, s, rst, nep, m, elist);
  vc_launched("k_vc_empty launch");
  // 2. streams over the segments; the refusal of the host builder
  int32_t* rb = (int32_t*)M.temp((size_t)(S + 1) * 4);
  hipLaunchKernelGGL(k_vc_parts, dim3(vc_grid(S + 1)), dim3(kWG), 0, s, vrst, nv, nnz, S, rb);
  vc_launched("k_vc_parts launch");
  int64_t* bound = (int64_t*)M.temp((size_t)S * 16);
  hipLaunchKernelGGL(k_vc_bound, dim3(vc_wave_grid(S)), dim3(kWG), 0, s, vrst, rb, S, bound);
  vc_launched("k_vc_bound launch");
  const int64_t slots_bound = vc_scan(M, bound, bound + S, S, s);
  if (slots_bound > 16 * nnz + (int64_t(1) << 20) || slots_bound > 2147483647LL) {
    set_last_error("CVR: too many streams for the entries (at least " +
                   std::to_string(slots_bound) + " packed slots for " + std::to_string(nnz) +
                   " entries)");
    return RVVLA_ERR_UNSUPPORTED_SPMV;
  }
  // 3. pass 1: packs and non-final records per stream
  int64_t* cnt = (int64_t*)M.temp((size_t)S * 32);
  int64_t *npk = cnt, *nrc = cnt + S, *ps = cnt + 2 * S, *rs = cnt + 3 * S;
  hipLaunchKernelGGL((k_vc_pack<V, V, 0, false>), dim3(vc_wave_grid(S)), dim3(kWG), 0, s, vrst,
                     vrow, col, val, rb, S, npk, nrc, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr);
  vc_launched("k_vc_pack launch");
  const int64_t npacks = vc_scan(M, npk, ps, S, s);
  if (npacks * 64 > 2147483647LL) {
    set_last_error("CVR: more than 2^31 packed entries");
    return RVVLA_ERR_LENGTH;
  }
  const int64_t nrec = vc_scan(M, nrc, rs, S, s);
  int64_t* pack_start = (int64_t*)M.put(nullptr, (size_t)(S + 1) * 8);
  int32_t* rec_start = (int32_t*)M.put(nullptr, (size_t)S * 4);
  hipLaunchKernelGGL(k_vc_starts, dim3(vc_grid(S + 1)), dim3(kWG), 0, s, ps, rs, S, npacks,
                     pack_start, rec_start);
  vc_launched("k_vc_starts launch");
  //    pass 2: the packs, flush masks, records and final-flush rows
  uint64_t* flush = (uint64_t*)M.put(nullptr, (size_t)npacks * 8);
  int32_t* rec_rows = (int32_t*)M.put(nullptr, (size_t)nrec * 4);
  int32_t* tails = (int32_t*)M.put(nullptr, (size_t)S * 64 * 4);
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint32_t* ocol = (uint32_t*)M.put(nullptr, (size_t)npacks * 64 * 4);
  void* oval = M.put(nullptr, (size_t)npacks * 64 * vbytes);
  if (M.dtype == RVVLA_BF16)
    vc_write<V, uint16_t, 1>(vrst, vrow, col, val, rb, S, ps, rs, flush, rec_rows, tails, ocol,
                             oval, s);
  else if (M.dtype == RVVLA_F16)
    vc_write<V, uint16_t, 2>(vrst, vrow, col, val, rb, S, ps, rs, flush, rec_rows, tails, ocol,
                             oval, s);
  else
    vc_write<V, V, 0>(vrst, vrow, col, val, rb, S, ps, rs, flush, rec_rows, tails, ocol, oval, s);
  // 4. cut rows: slots of their flushes, and their list
  int64_t* sl = (int64_t*)M.temp((size_t)m * 48 + 8);
  int64_t *longf = sl + m, *shortf = sl + 2 * m, *sbase = sl + 3 * m, *li = sl + 4 * m,
          *si = sl + 5 * m;
  hipLaunchKernelGGL(k_vc_cut_flags, dim3(vc_grid(m)), dim3(kWG), 0, s, rst, m, cap, sl, longf,
                     shortf);
  vc_launched("k_vc_cut_flags launch");
  const int64_t nslots = vc_scan(M, sl, sbase, m, s);
  const int64_t nlong = vc_scan(M, longf, li, m, s);
  const int64_t nshort = vc_scan(M, shortf, si, m, s);
  if (nslots > 2147483647LL) {
    set_last_error("CVR: too many partial sums");
    return RVVLA_ERR_LENGTH;
  }
  int32_t* split_rows = nullptr;
  int32_t* split_ptr = nullptr;
  void* part = nullptr;
  if (nslots > 0) {
    const int64_t N = nrec + 64 * S;
    uint32_t* key = (uint32_t*)M.temp((size_t)N * 16);
    uint32_t* gidx = key + N;
    uint32_t* skey = key + 2 * N;
    uint32_t* sg = key + 3 * N;
    int64_t* loc = (int64_t*)M.temp((size_t)N * 8);
    hipLaunchKernelGGL(k_vc_flush_keys, dim3(vc_grid(N)), dim3(kWG), 0, s, rst, rs, S, nrec, cap,
                       rec_rows, tails, key, gidx, loc);
    vc_launched("k_vc_flush_keys launch");
    size_t tmp = 0;
    int rc = hip_check(rocprim::radix_sort_pairs(nullptr, tmp, key, skey, gidx, sg, (size_t)N, 0, 32,
                                                 s),
                       "radix sort size");
    if (rc) return rc;
    void* tbuf = M.temp(tmp);
    rc = hip_check(rocprim::radix_sort_pairs(tbuf, tmp, key, skey, gidx, sg, (size_t)N, 0, 32, s),
                   "radix sort");
    if (rc) return rc;
    hipLaunchKernelGGL(k_vc_flush_slots, dim3(vc_grid(N)), dim3(kWG), 0, s, skey, sg, loc, N,
                       rec_rows, tails);
    vc_launched("k_vc_flush_slots launch");
    split_rows = (int32_t*)M.put(nullptr, (size_t)(nlong + nshort) * 4);
    split_ptr = (int32_t*)M.put(nullptr, (size_t)(nlong + nshort) * 8);
    hipLaunchKernelGGL(k_vc_cut_list, dim3(vc_grid(m)), dim3(kWG), 0, s, rst, m, cap, sbase, li, si,
                       nlong, split_rows, split_ptr);
    vc_launched("k_vc_cut_list launch");
    part = M.put(nullptr, (size_t)nslots * (M.dtype == RVVLA_F64 ? 8 : 4));
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device CVR build")) return rc;
  M.free_temps();
  d.nstreams = S;
  d.nempty = m - nne;
  d.npacks = npacks;
  d.nrec = nrec;
  d.pack_start = pack_start;
  d.rec_start = rec_start;
  d.flush = flush;
  d.rec_rows = rec_rows;
  d.tail_rows = tails;
  d.col = ocol;
  d.val = oval;
  d.empty_rows = elist;
  d.nsplit = nlong + nshort;
  d.nsplit_long = nlong;
  d.nslots = nslots;
  d.split_rows = split_rows;
  d.split_ptr = split_ptr;
  d.part = part;
  return RVVLA_OK;
}

template int build_cvr_device<double>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const double*, int64_t, hipStream_t);
template int build_cvr_device<float>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                     const int32_t*, const float*, int64_t, hipStream_t);

}  // namespace rvvla
