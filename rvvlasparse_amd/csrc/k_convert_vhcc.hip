// k_convert_vhcc.hip -- device-side CSR -> VHCC (vec_size = 64) builder.
//
// The device image that convert(RVVLA_FMT_VHCC) + upload give (the layout
// of convert_vhcc.cpp after Tang et al., CGO 2015, re-laid for the kernel),
// built from a CSR already on the device:
//  1. per entry its row (binary search) and whether it ends the row;
//  2. per partition of (nearly) equal vector counts: a forced row end on
//     its last entry when that row goes on past it, and, when its first
//     entry continues a row, the redirect of its first row end to the
//     partition's overflow slot (m + t); the hand-over flags of the fix-up;
//  3. a scan of the row-end flags orders the row ends in stream order
//     (row_arr, and each partition's first one, rbase); one wavefront per
//     vector ballots its row-end mask;
//  4. the rows that receive no row end (empty rows) by flags and a scan;
//     the entries padded to whole vectors.
#include <rocprim/device/device_scan.hpp>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

__global__ void k_vh_ends(const int32_t* __restrict__ rst, int64_t m, int64_t nnz,
                          int32_t* __restrict__ erow, uint8_t* __restrict__ isend,
                          int32_t* __restrict__ tgt) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e >= nnz) return;
  int64_t lo = 0, hi = m;  // the row of e: the last r with rst[r] <= e
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rst[mid] <= e) lo = mid;
    else hi = mid;
  }
  erow[e] = (int32_t)lo;
  const bool end = e == (int64_t)rst[lo + 1] - 1;
  isend[e] = end ? 1 : 0;
  tgt[e] = (int32_t)lo;
}

// partition t: forced end, overflow redirect, continuation (cont / last_row)
__global__ void k_vh_parts(const int32_t* __restrict__ erow, int64_t m, int64_t nnz, int64_t nvec,
                           int64_t T, int32_t* __restrict__ sv, uint8_t* __restrict__ isend,
                           int32_t* __restrict__ tgt, uint8_t* __restrict__ cont,
                           int32_t* __restrict__ last_row) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t > T) return;
  sv[t] = (int32_t)(nvec * t / T);
  if (t == T) return;
  const int64_t s0 = nvec * t / T, s1 = nvec * (t + 1) / T;
  const int64_t a = s0 * 64 < nnz ? s0 * 64 : nnz, b = s1 * 64 < nnz ? s1 * 64 : nnz;
  cont[t] = 0;
  last_row[t] = -1;
  if (a >= b) return;
  if (!isend[b - 1]) isend[b - 1] = 2;  // forced end: the row goes on past the partition
  if (t > 0 && a > 0 && erow[a] == erow[a - 1]) {
    for (int64_t e = a; e < b; ++e)
      if (isend[e]) {
        tgt[e] = (int32_t)(m + t);
        break;
      }
    cont[t] = 1;
    last_row[t] = erow[a];
  }
}

__global__ void k_vh_flags(const uint8_t* __restrict__ isend, int64_t nnz, int64_t* __restrict__ f) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e < nnz) f[e] = isend[e] ? 1 : 0;
}

__global__ void k_vh_row_arr(const uint8_t* __restrict__ isend, const int32_t* __restrict__ tgt,
                             const int64_t* __restrict__ epos, int64_t nnz, int64_t m,
                             uint32_t* __restrict__ row_arr, int64_t* __restrict__ hit) {
  const int64_t e = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (e >= nnz || !isend[e]) return;
  row_arr[epos[e]] = (uint32_t)tgt[e];
  if (tgt[e] < m) hit[tgt[e]] = 0;  // hit starts at 1 (= no row end yet)
}

__global__ __launch_bounds__(kWG) void k_vh_eor(const uint8_t* __restrict__ isend, int64_t nnz,
                                                int64_t nvec, uint64_t* __restrict__ eor) {
  const int64_t v = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  if (v >= nvec) return;
  const int64_t e = v * 64 + (threadIdx.x & 63);
  const uint64_t w = __ballot(e < nnz && isend[e] != 0);
  if ((threadIdx.x & 63) == 0) eor[v] = w;
}

__global__ void k_vh_fixup_info(const int64_t* __restrict__ epos, const int32_t* __restrict__ sv,
                                const uint8_t* __restrict__ cont,
                                const int32_t* __restrict__ last_row, int64_t T, int64_t nnz,
                                int64_t nends, int32_t* __restrict__ rbase,
                                int32_t* __restrict__ fwd, uint8_t* __restrict__ pass) {
  const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (t >= T) return;
  const int64_t a = (int64_t)sv[t] * 64;
  rbase[t] = (int32_t)(a < nnz ? epos[a] : nends);
  const bool nxt = t + 1 < T && cont[t + 1];
  const int32_t nrow = nxt ? last_row[t + 1] : -1;
  const bool p = cont[t] && nxt && last_row[t] == nrow;
  pass[t] = p ? 1 : 0;
  fwd[t] = (nxt && !p) ? nrow : -1;
}

__global__ void k_vh_empty_list(const int64_t* __restrict__ hit, const int64_t* __restrict__ hpos,
                                int64_t m, int32_t* __restrict__ list) {
  const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (r < m && hit[r]) list[hpos[r]] = (int32_t)r;
}

__global__ void k_vh_fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class V, class S, int NARROW>
__global__ void k_vh_entries(const int32_t* __restrict__ col, const V* __restrict__ val,
                             int64_t nnz, int64_t slots, uint32_t* __restrict__ ocol,
                             S* __restrict__ oval) {
  const int64_t j = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (j >= slots) return;
  const bool in = j < nnz;
  ocol[j] = in ? (uint32_t)col[j] : 0u;
  const V v = in ? val[j] : V(0);
  if constexpr (NARROW == 1) oval[j] = dev_bf16((float)v);
  else if constexpr (NARROW == 2) oval[j] = dev_f16((float)v);
  else oval[j] = v;
}

namespace {

unsigned vh_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWG - 1) / kWG); }

void vh_launched(const char* what) {
  if (int rc = hip_check(hipGetLastError(), what)) throw Error(rc, last_error());
}

int64_t vh_scan(rvvla_mat_s& M, const int64_t* cnt, int64_t* out, int64_t n, hipStream_t s) {
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

}  // namespace

template <class V>
int build_vhcc_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t nparts, hipStream_t s) {
  DevVHCC& d = M.vhcc;
  const int64_t nvec = (nnz + 63) / 64;
  const int64_t T = std::max<int64_t>(1, std::min<int64_t>(nparts, std::max<int64_t>(nvec, 1)));
  int32_t* erow = (int32_t*)M.temp((size_t)nnz * 9 + 8);
  int32_t* tgt = erow + nnz;
  uint8_t* isend = (uint8_t*)(tgt + nnz);
  if (nnz > 0) {
    hipLaunchKernelGGL(k_vh_ends, dim3(vh_grid(nnz)), dim3(kWG), 0, s, rst, m, nnz, erow, isend,
                       tgt);
    vh_launched("k_vh_ends launch");
  }
  int32_t* sv = (int32_t*)M.put(nullptr, (size_t)(T + 1) * 4);
  uint8_t* cont = (uint8_t*)M.temp((size_t)T + 8);
  int32_t* last_row = (int32_t*)M.temp((size_t)T * 4 + 8);
  hipLaunchKernelGGL(k_vh_parts, dim3(vh_grid(T + 1)), dim3(kWG), 0, s, erow, m, nnz, nvec, T, sv,
                     isend, tgt, cont, last_row);
  vh_launched("k_vh_parts launch");
  int64_t* f = (int64_t*)M.temp((size_t)nnz * 16 + 8);
  int64_t* epos = f + nnz;
  if (nnz > 0) {
    hipLaunchKernelGGL(k_vh_flags, dim3(vh_grid(nnz)), dim3(kWG), 0, s, isend, nnz, f);
    vh_launched("k_vh_flags launch");
  }
  const int64_t nends = vh_scan(M, f, epos, nnz, s);
  uint32_t* row_arr = (uint32_t*)M.put(nullptr, (size_t)nends * 4);
  int64_t* hit = (int64_t*)M.temp((size_t)m * 16 + 8);
  int64_t* hpos = hit + m;
  hipLaunchKernelGGL(k_vh_fill_i64, dim3(vh_grid(m)), dim3(kWG), 0, s, hit, m, (int64_t)1);
  vh_launched("k_vh_fill_i64 launch");
  if (nnz > 0) {
    hipLaunchKernelGGL(k_vh_row_arr, dim3(vh_grid(nnz)), dim3(kWG), 0, s, isend, tgt, epos, nnz, m,
                       row_arr, hit);
    vh_launched("k_vh_row_arr launch");
  }
  uint64_t* eor = (uint64_t*)M.put(nullptr, (size_t)nvec * 8);
  if (nvec > 0) {
    hipLaunchKernelGGL(k_vh_eor, dim3((unsigned)((nvec + kWG / 64 - 1) / (kWG / 64))), dim3(kWG),
                       0, s, isend, nnz, nvec, eor);
    vh_launched("k_vh_eor launch");
  }
  int32_t* rbase = (int32_t*)M.put(nullptr, (size_t)T * 4);
  int32_t* fwd = (int32_t*)M.put(nullptr, (size_t)T * 4);
  uint8_t* pass = (uint8_t*)M.put(nullptr, (size_t)T);
  hipLaunchKernelGGL(k_vh_fixup_info, dim3(vh_grid(T)), dim3(kWG), 0, s, epos, sv, cont, last_row,
                     T, nnz, nends, rbase, fwd, pass);
  vh_launched("k_vh_fixup_info launch");
  const int64_t nempty = vh_scan(M, hit, hpos, m, s);
  int32_t* elist = (int32_t*)M.put(nullptr, (size_t)nempty * 4);
  hipLaunchKernelGGL(k_vh_empty_list, dim3(vh_grid(m)), dim3(kWG), 0, s, hit, hpos, m, elist);
  vh_launched("k_vh_empty_list launch");
  const size_t vbytes = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  uint32_t* ocol = (uint32_t*)M.put(nullptr, (size_t)(nvec * 64) * 4);
  void* oval = M.put(nullptr, (size_t)(nvec * 64) * vbytes);
  if (nvec > 0) {
    const unsigned g = vh_grid(nvec * 64);
    if (M.dtype == RVVLA_BF16)
      hipLaunchKernelGGL((k_vh_entries<V, uint16_t, 1>), dim3(g), dim3(kWG), 0, s, col, val, nnz,
                         nvec * 64, ocol, (uint16_t*)oval);
    else if (M.dtype == RVVLA_F16)
      hipLaunchKernelGGL((k_vh_entries<V, uint16_t, 2>), dim3(g), dim3(kWG), 0, s, col, val, nnz,
                         nvec * 64, ocol, (uint16_t*)oval);
    else
      hipLaunchKernelGGL((k_vh_entries<V, V, 0>), dim3(g), dim3(kWG), 0, s, col, val, nnz,
                         nvec * 64, ocol, (V*)oval);
    vh_launched("k_vh_entries launch");
  }
  if (int rc = hip_check(hipStreamSynchronize(s), "device VHCC build")) return rc;
  M.free_temps();
  const size_t acc = M.dtype == RVVLA_F64 ? 8 : 4;
  d.nparts = T;
  d.nempty = nempty;
  d.nvec = nvec;
  d.nends = nends;
  d.start_vec = sv;
  d.rbase = rbase;
  d.eor = eor;
  d.row_arr = row_arr;
  d.fwd_row = fwd;
  d.pass = pass;
  d.col = ocol;
  d.val = oval;
  d.empty_rows = elist;
  d.scratch = M.put(nullptr, (size_t)T * 2 * acc + 16);
  M.device_bytes -= (int64_t)((size_t)T * 2 * acc + 16);
  return RVVLA_OK;
}

template int build_vhcc_device<double>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                       const int32_t*, const double*, int64_t, hipStream_t);
template int build_vhcc_device<float>(rvvla_mat_s&, int64_t, int64_t, const int32_t*,
                                      const int32_t*, const float*, int64_t, hipStream_t);

}  // namespace rvvla
