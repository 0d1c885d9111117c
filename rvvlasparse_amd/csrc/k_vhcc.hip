// k_vhcc.hip -- VHCC SpMV for gfx950 (vec_size = 64: one vector per
// wavefront step, one partition per wavefront).
//
// Semantics follow the reference vector kernel (reference
// src/rvv/spmv_vhcc.cpp:73-162 and its epilogue :281-313): vectors of the
// CSR-ordered entry stream are multiplied lane-wise; at a vector holding row
// ends, a segmented inclusive scan gives every ending row its sum, which is
// written to row_arr's row; the open segment carries into lane 0 of the
// next vector; rows crossing partitions are completed by a fix-up.
// gfx950 form:
//  * vectors without a row end add into lane-private sums (no cross-lane
//    work); the next row-end vector folds them into its carry;
//  * the segmented scan's segments follow from the row-end mask alone (a
//    lane's segment starts after the highest row end below it), so only the
//    64-bit row-end mask per vector is read; the scan runs on DPP row
//    shifts inside each row of 16 lanes plus three row carries (no LDS
//    permutes);
//  * with 32-bit accumulation, each XCD walks a contiguous run of
//    partitions (xcd_run_index);
//  * instead of the reference's overflow rows and atomic fix-up, a partition
//    stores the partial of the row it continues (first_part) and of the row
//    it hands on (fwd_part); one owner thread per handed-on row adds the
//    pieces in partition order (deterministic), y = alpha*sum + beta*y.
#include "kcommon.h"

namespace rvvla {

struct VhccArgs {
  const int32_t* start_vec;  // nparts + 1
  const int32_t* rbase;
  const uint64_t* eor;
  const uint32_t* row_arr;
  const int32_t* fwd_row;    // row handed on to the next partition, or -1
  const uint8_t* pass;       // partition's only row end is the handed-on one
  const uint32_t* col;
  const int32_t* empty_rows;
  int64_t nparts, nempty, m;
};

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_vhcc(VhccArgs a, const VT* __restrict__ val,
                                               const typename Val<VT>::io* __restrict__ x,
                                               typename Val<VT>::io* __restrict__ y,
                                               typename Val<VT>::acc* __restrict__ first_part,
                                               typename Val<VT>::acc* __restrict__ fwd_part,
                                               typename Val<VT>::acc alpha,
                                               typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  using Ac = Acc<A, COMP>;
  const int64_t t = __builtin_amdgcn_readfirstlane((int)((sizeof(A) < 8 ? xcd_run_index((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x) * (kWG / 64) + (threadIdx.x >> 6)));
  if (t >= a.nparts) return;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  const int64_t v0 = a.start_vec[t], v1 = a.start_vec[t + 1];
  int64_t rank = a.rbase[t];
  const bool has_fwd = a.fwd_row[t] >= 0;
  const int64_t last_entry = v1 * 64 - 1;
  const uint32_t* __restrict__ colp = a.col;
  const uint32_t* __restrict__ rowp = a.row_arr;
  const uint64_t* __restrict__ eorp = a.eor;
  Ac hist, carry;
  bool pending = false;  // hist holds sums of vectors without row ends
  // kBatch vectors at a time: all their loads (columns, values, masks, the
  // ending rows) and gathers are issued before the scans
  constexpr int kBatch = 4;
  for (int64_t vb = v0; vb < v1; vb += kBatch) {
    uint32_t c[kBatch];
    A vv[kBatch], p[kBatch];
    uint64_t mk[kBatch];
    uint32_t rr[kBatch];
    if (vb + kBatch <= v1) {  // a whole batch: straight-line loads
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int64_t e = (vb + u) * 64 + lane;
        c[u] = colp[e];
        vv[u] = Val<VT>::get(val, e);
        mk[u] = eorp[vb + u];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const bool in = vb + u < v1;
        const int64_t e = (vb + u) * 64 + lane;
        c[u] = in ? colp[e] : 0u;
        vv[u] = in ? Val<VT>::get(val, e) : A(0);
        mk[u] = in ? eorp[vb + u] : 0ull;
      }
    }
    int64_t rk = rank;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      p[u] = vv[u] * (A)x[c[u]];
      rr[u] = ((mk[u] >> lane) & 1u) ? rowp[rk + __popcll(mk[u] & below)] : 0u;
      rk += __popcll(mk[u]);
    }
    rank = rk;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (vb + u >= v1) break;
      const uint64_t m = mk[u];
      if (!m) {
        hist.add(p[u]);
        pending = true;
        continue;
      }
      if (pending) {
        const Ac h = group_reduce(hist, 64);
        carry.merge(h);
        hist = Ac();
        pending = false;
      }
      Ac sum;
      sum.add(p[u]);
      if (lane == 0) sum.merge(carry);
      sum = wave_seg_scan(sum, 1u | (m << 1), lane);
      if ((m >> lane) & 1u) {
        const uint32_t r = rr[u];
        const int64_t e = (vb + u) * 64 + lane;
        if (r >= (uint32_t)a.m) {
          first_part[t] = sum.value();  // the row this partition continues
        } else if (has_fwd && e == last_entry) {
          fwd_part[t] = sum.value();    // the row handed on to the next partition
        } else {
          y[r] = alpha * sum.value() + scaled_y(beta, y, r);
        }
      }
      carry = (m >> 63) ? Ac() : acc_readlane(sum, 63);
    }
  }
}

template <class VT, bool COMP>
__global__ __launch_bounds__(kWG) void k_vhcc_fixup(VhccArgs a,
                                                     typename Val<VT>::io* __restrict__ y,
                                                     const typename Val<VT>::acc* __restrict__ first_part,
                                                     const typename Val<VT>::acc* __restrict__ fwd_part,
                                                     typename Val<VT>::acc alpha,
                                                     typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (i < a.nparts) {
    const int32_t r = a.fwd_row[i];
    if (r < 0) return;
    Acc<A, COMP> s;
    s.add(fwd_part[i]);
    for (int64_t q = i + 1; q < a.nparts; ++q) {
      s.add(first_part[q]);
      if (!a.pass[q]) break;
    }
    y[r] = alpha * s.value() + scaled_y(beta, y, r);
  } else if (i < a.nparts + a.nempty) {
    const int32_t r = a.empty_rows[i - a.nparts];
    y[r] = scaled_y(beta, y, r);
  }
}

template <class VT>
static int run_vhcc(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                    hipStream_t s, void* scratch) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const DevVHCC& d = A.vhcc;
  VhccArgs a{d.start_vec, d.rbase, d.eor, d.row_arr, d.fwd_row, d.pass, d.col, d.empty_rows,
             d.nparts, d.nempty, A.m};
  constexpr bool COMP = sizeof(VT) == 2;  // 16-bit values: TwoSum throughout
  AC* fp = (AC*)(scratch ? scratch : d.scratch);
  AC* wp = fp + d.nparts;
  if (d.nparts > 0) {
    const unsigned g0 = (unsigned)((d.nparts + kWG / 64 - 1) / (kWG / 64));
    const unsigned g = sizeof(AC) < 8 ? (unsigned)xcd_grid(g0) : g0;
    hipLaunchKernelGGL((k_vhcc<VT, COMP>), dim3(g), dim3(kWG), 0, s, a, (const VT*)d.val, (const IO*)x,
                       (IO*)y, fp, wp, (AC)alpha, (AC)beta);
    if (int rc = hip_check(hipGetLastError(), "k_vhcc launch")) return rc;
  }
  const int64_t nfix = d.nparts + d.nempty;
  if (nfix > 0) {
    hipLaunchKernelGGL((k_vhcc_fixup<VT, COMP>), dim3((unsigned)((nfix + kWG - 1) / kWG)), dim3(kWG), 0,
                       s, a, (IO*)y, fp, wp, (AC)alpha, (AC)beta);
    return hip_check(hipGetLastError(), "k_vhcc_fixup launch");
  }
  return RVVLA_OK;
}

int launch_vhcc(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch) {
  switch (A.dtype) {
    case RVVLA_F64: return run_vhcc<double>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_vhcc<float>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_vhcc<bf16_t>(A, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_vhcc<_Float16>(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
