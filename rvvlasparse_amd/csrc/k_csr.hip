// k_csr.hip -- CSR SpMV / SpMM for gfx950.
//
// Semantics: y_i = alpha * sum_j Val_j * x[Col_j] + beta * y_i, as the
// reference scalar CSR (reference src/scalar/spmv_mtx_crs.cpp:37-43), except
// that y is not read when beta == 0.
//
// One workgroup (256 threads = 4 wavefronts) per row block, built once at
// upload:
//  * stream block: up to 256 consecutive rows with <= 2048 entries in all.
//    The entries are read fully coalesced, multiplied by x and staged in
//    LDS; then a row of at most kShortSum entries is summed out of LDS by
//    its own thread in entry order, a longer one by a whole wavefront
//    (strided lanes, then a butterfly).
//  * long block: one row with more than 1024 entries; all 256 threads
//    stride over it, then wave and workgroup reductions.  In the 16-bit
//    value forms the partial sums are compensated (TwoSum) there.
//
// Blocks are (first, end) row pairs, long-row blocks first (longest first)
// so that no long row starts at the end of the launch.
//
// Column-sliced image (csr_xslice.cpp): the same kernel runs over the
// compressed (row, slice) rows, each block on the XCD of its slice so that
// each XCD gathers from its own slice of x (entries and columns are then streamed
// with nontemporal loads, to leave L2 to x), and writes one partial per pair
// (alpha = 1, beta = 0); k_xs_combine then adds each row's partials in
// slice order and applies alpha and beta.
#include <type_traits>

#include "kcommon.h"

namespace rvvla {

constexpr int kShortSum = 16;

template <class VT, bool COMP, bool NT>
__global__ __launch_bounds__(kWG) void k_csr_adaptive(
    const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const int32_t* __restrict__ blocks,
    const typename Val<VT>::io* __restrict__ x, typename Val<VT>::io* __restrict__ y,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  __shared__ A prod[kStreamCap];
  __shared__ A wsum[2 * (kWG / 64)];
  const int tid = threadIdx.x;
  const int r0 = blocks[2 * blockIdx.x], r1 = blocks[2 * blockIdx.x + 1];
  if (r0 == r1) return;  // padding block of a sliced image (uniform)
  const int e0 = rst[r0], e1 = rst[r1];
  const int nnz = e1 - e0, nrows = r1 - r0;

  if (nrows == 1 && nnz > kLongRow) {
    Acc<A, COMP> a;
    int e = e0 + tid;
    for (; e + 3 * kWG < e1; e += 4 * kWG) {
      const int c0 = ldi<NT>(col, e), c1 = ldi<NT>(col, e + kWG);
      const int c2 = ldi<NT>(col, e + 2 * kWG), c3 = ldi<NT>(col, e + 3 * kWG);
      const A v0 = ldv<NT>(val, e), v1 = ldv<NT>(val, e + kWG);
      const A v2 = ldv<NT>(val, e + 2 * kWG), v3 = ldv<NT>(val, e + 3 * kWG);
      a.add(v0 * (A)x[c0]);
      a.add(v1 * (A)x[c1]);
      a.add(v2 * (A)x[c2]);
      a.add(v3 * (A)x[c3]);
    }
    for (; e < e1; e += kWG) a.add(ldv<NT>(val, e) * (A)x[ldi<NT>(col, e)]);
    a = group_reduce(a, 64);
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
      wsum[w] = a.s;
      wsum[kWG / 64 + w] = a.c;
    }
    __syncthreads();
    if (tid == 0) {
      Acc<A, COMP> t;
      t.s = wsum[0];
      t.c = wsum[kWG / 64];
      for (int i = 1; i < kWG / 64; ++i) {
        Acc<A, COMP> o;
        o.s = wsum[i];
        o.c = wsum[kWG / 64 + i];
        t.merge(o);
      }
      y[r0] = alpha * t.value() + scaled_y(beta, y, r0);
    }
    return;
  }

  // stream block: row bounds of thread tid's row (fetched before the
  // barrier), then coalesced products into LDS
  const int t_row = r0 + tid;
  int lo = 0, hi = 0;
  if (tid < nrows) {
    lo = rst[t_row] - e0;
    hi = rst[t_row + 1] - e0;
  }
  {
    int k = tid;
    for (; k + 3 * kWG < nnz; k += 4 * kWG) {
      const int c0 = ldi<NT>(col, e0 + k), c1 = ldi<NT>(col, e0 + k + kWG);
      const int c2 = ldi<NT>(col, e0 + k + 2 * kWG), c3 = ldi<NT>(col, e0 + k + 3 * kWG);
      const A v0 = ldv<NT>(val, e0 + k), v1 = ldv<NT>(val, e0 + k + kWG);
      const A v2 = ldv<NT>(val, e0 + k + 2 * kWG), v3 = ldv<NT>(val, e0 + k + 3 * kWG);
      prod[k] = v0 * (A)x[c0];
      prod[k + kWG] = v1 * (A)x[c1];
      prod[k + 2 * kWG] = v2 * (A)x[c2];
      prod[k + 3 * kWG] = v3 * (A)x[c3];
    }
    for (; k < nnz; k += kWG) prod[k] = ldv<NT>(val, e0 + k) * (A)x[ldi<NT>(col, e0 + k)];
  }
  __syncthreads();
  // rows of at most kShortSum products: their own thread, in entry order;
  // longer rows: the whole wavefront of the row's thread, one row at a time
  const bool is_long = tid < nrows && hi - lo > kShortSum;
  if (tid < nrows && !is_long) {
    Acc<A, COMP> s;
    for (int k = lo; k < hi; ++k) s.add(prod[k]);
    y[t_row] = alpha * s.value() + scaled_y(beta, y, t_row);
  }
  const int lane = tid & 63;
  uint64_t lm = __ballot(is_long);
  while (lm) {
    const int l = __ffsll((unsigned long long)lm) - 1;
    lm &= lm - 1;
    const int rlo = __shfl(lo, l, 64), rhi = __shfl(hi, l, 64);
    Acc<A, COMP> a;
    for (int k = rlo + lane; k < rhi; k += 64) a.add(prod[k]);
    a = group_reduce(a, 64);
    if (lane == 0) {
      const int r = r0 + (tid - lane) + l;
      y[r] = alpha * a.value() + scaled_y(beta, y, r);
    }
  }
}

// y[r] = alpha * (sum over slices p in order of part[pair(r, p)]) + beta*y[r].
// A wavefront covers kCombineGroups consecutive groups of 64 rows, one row
// of each per lane; the pair of row r (group q) in slice p is base[p] +
// pref[p*nq + q] + (lower lanes with bit p set).  All partial loads of all
// groups are issued before the (in-order) sums.
constexpr int kCombineGroups = 1;
template <class IO, class A, bool COMP>
__global__ __launch_bounds__(kWG) void k_xs_combine(
    const uint16_t* __restrict__ mask, const int32_t* __restrict__ base,
    const int32_t* __restrict__ pref, const IO* __restrict__ part, IO* __restrict__ y,
    int64_t m, int64_t nq, int P, A alpha, A beta) {
  constexpr int G = kCombineGroups;
  const int lane = threadIdx.x & 63;
  const int64_t q0 =
      ((int64_t)blockIdx.x * (kWG / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * G;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned mk[G];
  A v[G][16];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t r = (q0 + g) * 64 + lane;
    mk[g] = r < m ? mask[r] : 0u;
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const bool has = p < P && ((mk[g] >> p) & 1u);
      const uint64_t bal = __ballot(has);
      v[g][p] = A(0);
      if (has) v[g][p] = (A)part[base[p] + pref[(int64_t)p * nq + q0 + g] + __popcll(bal & below)];
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t r = (q0 + g) * 64 + lane;
    Acc<A, COMP> s;
#pragma unroll
    for (int p = 0; p < 16; ++p)
      if ((mk[g] >> p) & 1u) s.add(v[g][p]);
    if (r < m) y[r] = alpha * s.value() + scaled_y(beta, y, r);
  }
}

// ---- LDS-sliced image (csr_xlds.cpp) --------------------------------------
//
// Slice pass: one workgroup per CU (shape: XlShape) runs its units (slice p,
// tasks t0 .. t1-1) in turn: x[p W, p W + W) into LDS, then the tasks,
// wavefront w taking tasks t0 + w, t0 + w + 16, ...  In a task lane l holds
// entries l*8 .. l*8+7 (one 16-byte column load, 16-byte value loads).
// Bit 15 marks the first entry of a piece and every task starts a piece.
// A lane adds its entries in order; a piece that starts and ends inside
// the lane is written at once.  The entries before the lane's first mark
// (its head) belong to a piece opened in a lower lane, those from its last
// mark on (its tail) to a piece that may continue in higher lanes.  A
// segmented scan across the lanes (a segment starts at each lane holding a
// mark) gives the running sum of the open piece at every lane: a lane with
// marks completes the piece open below it (the scan of the lane below plus
// its head), the last lane completes the task's last piece.  A piece's slot
// is the count of marks before it, so a task fills consecutive slots.  The
// 16-bit forms add with TwoSum compensation throughout.
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Launch shape of the slice pass: FP64 (x slice of 16384 doubles) runs 8
// wavefronts a CU, each with 4 tasks in flight; the 32-bit-x forms run 16
// wavefronts with one task ahead.  Both stage a task's partials in LDS, so
// that they leave in whole lines (writing them from the lanes straight to
// their slots, with 16 FP64 wavefronts a CU, measured slower).
template <class VT>
struct XlShape {
  static constexpr bool F64 = sizeof(VT) == 8;
  static constexpr int WG = F64 ? 512 : 1024;
  static constexpr int AHEAD = F64 ? 4 : 1;
  // entries streamed with nontemporal loads (measured faster for the
  // 32-bit-x forms, slower for FP64)
  static constexpr bool NT = !F64;
};

__device__ __forceinline__ uint4 xl_ld_nt(const uint4* p) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 t = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(t.x, t.y, t.z, t.w);
}

// A task's data as loaded: its descriptor, 8 columns and 8 values per lane
// (the values as raw 16-byte words).
template <class VT>
struct XlTaskData {
  static constexpr int NV = sizeof(VT) / 2;  // 16-byte words of 8 values
  int4 tk;
  uint4 cw;
  uint4 vw[NV];
};

template <class VT>
__device__ __forceinline__ void xl_fetch(XlTaskData<VT>& d, const uint16_t* __restrict__ col,
                                         const VT* __restrict__ val, const int4* __restrict__ task,
                                         int t, int lane) {
  d.tk = task[t];
  const uint4* cp = reinterpret_cast<const uint4*>(col + d.tk.x + lane * 8);
  // the task's values are stored word-interleaved: word i of lane l at
  // 16-byte word i * 64 + l of the task
  const uint4* vp = reinterpret_cast<const uint4*>(val + d.tk.x) + lane;
  if constexpr (XlShape<VT>::NT) {  // read once: nontemporal
    d.cw = xl_ld_nt(cp);
#pragma unroll
    for (int i = 0; i < XlTaskData<VT>::NV; ++i) d.vw[i] = xl_ld_nt(vp + i * 64);
  } else {
    d.cw = *cp;
#pragma unroll
    for (int i = 0; i < XlTaskData<VT>::NV; ++i) d.vw[i] = vp[i * 64];
  }
}

template <class VT>
__device__ __forceinline__ void xl_unpack(const XlTaskData<VT>& d, typename Val<VT>::acc (&v)[8]) {
  if constexpr (sizeof(VT) == 8) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = __builtin_bit_cast(double, ((uint64_t)d.vw[i].y << 32) | d.vw[i].x);
      v[2 * i + 1] = __builtin_bit_cast(double, ((uint64_t)d.vw[i].w << 32) | d.vw[i].z);
    }
  } else if constexpr (sizeof(VT) == 4) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      v[4 * i] = __uint_as_float(d.vw[i].x);
      v[4 * i + 1] = __uint_as_float(d.vw[i].y);
      v[4 * i + 2] = __uint_as_float(d.vw[i].z);
      v[4 * i + 3] = __uint_as_float(d.vw[i].w);
    }
  } else {
    const uint32_t w[4] = {d.vw[0].x, d.vw[0].y, d.vw[0].z, d.vw[0].w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if constexpr (std::is_same<VT, bf16_t>::value) {
        v[2 * i] = __uint_as_float(w[i] << 16);
        v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
      } else {
        v[2 * i] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[i] & 0xFFFFu));
        v[2 * i + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[i] >> 16));
      }
    }
  }
}

// Slots of a task staged in LDS: slot s of the task at s ^ ((s >> 5) & 7),
// so that neither the lanes' scattered writes nor the contiguous read back
// meet more than two-way bank conflicts.
__device__ __forceinline__ int xl_swz(int s) { return s ^ ((s >> 5) & 7); }

// Exclusive prefix count of v over the lanes (DPP row shifts, then the row
// totals), and the wavefront total.
__device__ __forceinline__ int xl_lane_prefix(int v, int lane, int& total) {
  int s = v;
  s += __builtin_amdgcn_update_dpp(0, s, 0x111, 0xF, 0xF, false);
  s += __builtin_amdgcn_update_dpp(0, s, 0x112, 0xF, 0xF, false);
  s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xF, false);
  s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xF, 0xF, false);
  const int r0 = __builtin_amdgcn_readlane(s, 15), r1 = __builtin_amdgcn_readlane(s, 31);
  const int r2 = __builtin_amdgcn_readlane(s, 47), r3 = __builtin_amdgcn_readlane(s, 63);
  const int row = lane >> 4;
  s += (row >= 1 ? r0 : 0) + (row >= 2 ? r1 : 0) + (row >= 3 ? r2 : 0);
  total = r0 + r1 + r2 + r3;
  return s - v;
}

// One task.  The lane's walk writes the running sum of its open piece to
// that piece's stage slot where the piece leaves the lane (the next entry
// opens another piece, or the lane's last entry), so each write is the
// lane's complete share of the piece; a piece that continues into higher
// lanes is written again, complete, by the lane that closes it (after
// every walk write, so it wins).  The stage then goes out in slot order.
template <class VT, bool COMP>
__device__ __forceinline__ void xl_task(const XlTaskData<VT>& d, const typename Val<VT>::io* xs,
                                        typename Val<VT>::acc* __restrict__ part,
                                        typename Val<VT>::acc* stg, int lane) {
  using A = typename Val<VT>::acc;
  using IO = typename Val<VT>::io;
  using Ac = Acc<A, COMP>;
  constexpr int XW = xl_slice_cols(sizeof(IO));
  const int ec = d.tk.y, pc0 = d.tk.z;
  A v[8];
  xl_unpack<VT>(d, v);
  const uint32_t cwd[4] = {d.cw.x, d.cw.y, d.cw.z, d.cw.w};
  uint32_t c[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c[2 * i] = cwd[i] & 0xFFFFu;
    c[2 * i + 1] = cwd[i] >> 16;
  }
  const int nv = ec - lane * 8;  // valid entries of this lane (may be <= 0)
  // the eight x gathers from LDS issued together: every index is inside
  // the slice, and the empty asm takes all eight values before any is
  // masked, so that none is moved into a branch of its own
  IO xv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) xv[k] = xs[c[k] & (XW - 1)];
  __asm__ volatile("" ::"v"(xv[0]), "v"(xv[1]), "v"(xv[2]), "v"(xv[3]), "v"(xv[4]), "v"(xv[5]),
                   "v"(xv[6]), "v"(xv[7]));
  A prod[8];
  bool fl[8];
  int nfl = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool ok = k < nv;
    fl[k] = ok && (c[k] & 0x8000u);
    nfl += fl[k] ? 1 : 0;
    prod[k] = ok ? v[k] * (A)xv[k] : A(0);
  }
  int tot;
  const int fb = xl_lane_prefix(nfl, lane, tot);  // marks in lower lanes, in the task
  Ac run, head;
  bool seen = false;
  int slot = fb - 1;  // task-relative slot of the lane's open piece
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool f = fl[k];
    const bool first = f && !seen;
    head.s = first ? run.s : head.s;
    head.c = first ? run.c : head.c;
    run.s = f ? A(0) : run.s;
    run.c = f ? A(0) : run.c;
    seen = seen || f;
    slot += f ? 1 : 0;
    run.add(prod[k]);
    // the piece's share leaves when the next entry opens another piece
    // (or at the lane's last entry)
    if (k == 7 || fl[k < 7 ? k + 1 : 7]) stg[xl_swz(slot)] = run.value();
  }
  // segmented inclusive scan of the open sums (lane 0 always holds a mark)
  const Ac S = wave_seg_scan(run, __ballot(seen), lane);
  Ac prev;
  prev.s = dpp_mov<0x138>(S.s);  // wave_shr:1: the lane below
  prev.c = dpp_mov<0x138>(S.c);
  prev.merge(head);
  if (seen && lane > 0) stg[xl_swz(fb - 1)] = prev.value();
  if (lane == 63) stg[xl_swz(tot - 1)] = S.value();
  // the stage back in slot order: whole-line stores
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int i = lane; i < tot; i += 64) part[pc0 + i] = stg[xl_swz(i)];
}

template <class VT, bool COMP>
__global__ __launch_bounds__(XlShape<VT>::WG) void k_xl_slices(
    const uint16_t* __restrict__ col, const VT* __restrict__ val, const int4* __restrict__ task,
    const int32_t* __restrict__ units, const int32_t* __restrict__ wg,
    const typename Val<VT>::io* __restrict__ x, int64_t n, typename Val<VT>::acc* __restrict__ part) {
  using IO = typename Val<VT>::io;
  using A = typename Val<VT>::acc;
  constexpr int WGS = XlShape<VT>::WG;
  constexpr int D = XlShape<VT>::AHEAD;
  constexpr int XW = xl_slice_cols(sizeof(IO));
  constexpr int XP = XW / WGS;
  constexpr int NW = WGS / 64;
  static_assert(XW % WGS == 0, "x slice per thread");
  static_assert(XW * sizeof(IO) + NW * kXlTask * sizeof(A) <= 163840, "LDS");
  __shared__ IO xs[XW];
  __shared__ A stage[NW * kXlTask];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  A* stg = stage + wave * kXlTask;
  const int u0 = wg[blockIdx.x], u1 = wg[blockIdx.x + 1];
  for (int u = u0; u < u1; ++u) {
    const int64_t p = units[3 * u];
    const int t0 = units[3 * u + 1], t1 = units[3 * u + 2];
    // the wavefront's first D tasks are fetched before the slice of x (a
    // task index past the unit fetches its last task again, unused)
    XlTaskData<VT> b[D];
    const int tw = t0 + wave;
#pragma unroll
    for (int i = 0; i < D; ++i) xl_fetch<VT>(b[i], col, val, task, min(tw + i * NW, t1 - 1), lane);
    __syncthreads();  // every wavefront is done with the previous slice
    {
      const int64_t c0 = p * XW;
      IO tmp[XP];
      if (n - c0 >= XW) {  // a whole slice: unguarded loads
#pragma unroll
        for (int j = 0; j < XP; ++j) tmp[j] = x[c0 + tid + j * WGS];
      } else {
        const int wl = (int)(n - c0);
#pragma unroll
        for (int j = 0; j < XP; ++j) {
          const int i = tid + j * WGS;
          tmp[j] = i < wl ? x[c0 + i] : IO(0);
        }
      }
#pragma unroll
      for (int j = 0; j < XP; ++j) xs[tid + j * WGS] = tmp[j];
    }
    __syncthreads();
    for (int t = tw; t < t1; t += D * NW) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        const int ti = t + i * NW;
        if (ti < t1) xl_task<VT, COMP>(b[i], xs, part, stg, lane);
        // the ring slot refilled D tasks ahead
        xl_fetch<VT>(b[i], col, val, task, min(ti + D * NW, t1 - 1), lane);
      }
    }
  }
}

// Combine of the LDS-sliced image, WG threads a block: block b (rows
// r0 .. r1-1, at most 2 WG of them) takes from each slice p its run of
// slots off[b S + p] .. off[(b+1) S + p] - 1.  The run lengths are scanned
// in LDS, numbering the block's partials 0 .. cnt-1 in slice order; each
// non-empty run marks its first partial in a bitmap and records itself as
// the run of every 64-partial word it begins.  Thread t takes partials
// t, t + WG, ...: a wavefront's 64 partials are one word, so a lane's run
// is the word's first run plus the marks up to it (a bit count), and its
// slot that run's offset from it; all loads are issued together.  Each
// partial goes to its row-major place lpos in LDS; then each row's
// partials are added in slice order (a thread per row, a wavefront for a
// row of more than 32 pieces) and alpha and beta applied.  The run bounds
// are loaded first and the rows' piece bounds with them.  A block of one
// row with more pieces than LDS holds adds its runs in place.  Consecutive
// blocks run on one XCD, so the lines their runs share stay in that XCD's
// L2.
template <class A, bool COMP, int WG>
__device__ __forceinline__ void xc_finish(Acc<A, COMP> a, A* wsum, int lane, int wave) {
  constexpr int NW = WG / 64;
  a = group_reduce(a, 64);
  if (lane == 0) {
    wsum[wave] = a.s;
    wsum[NW + wave] = a.c;
  }
}

template <class A, class IO, bool COMP, int WG>
__global__ __launch_bounds__(WG) void k_xl_combine(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cblk,
    const int32_t* __restrict__ off, const uint16_t* __restrict__ lpos,
    const A* __restrict__ part, IO* __restrict__ y, int64_t S, int nb, A alpha, A beta) {
  using Ac = Acc<A, COMP>;
  constexpr int CP = xl_comb_pieces(sizeof(A), WG);
  constexpr int NW = WG / 64;
  constexpr int PT = CP / WG;
  constexpr int RT = xl_comb_rows(WG) / WG;  // rows of a thread
  __shared__ A prod[CP];
  __shared__ uint32_t bits[CP / 32];  // run starts over the block's partials
  __shared__ int rw[CP / 64];         // run of each 64-partial word's first
  __shared__ int delta[WG];           // slot - partial index of each run
  __shared__ A wsum[2 * NW];
  __shared__ int wtot[NW];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int per = (nb + 7) / 8;
  const int b = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (b >= nb) return;
  const int32_t* o0 = off + (int64_t)b * S;
  const int32_t* o1 = o0 + S;
  // thread t: slice t's run (independent of the rows: loaded first, from
  // a clamped index so that no branch waits for it)
  const bool scan = S <= WG;
  const int pt = scan ? min(tid, (int)S - 1) : 0;
  const int a0 = o0[pt], a1 = o1[pt];
  const int r0 = cblk[2 * b], r1 = cblk[2 * b + 1];
  const int base = rowptr[r0], cnt = rowptr[r1] - base;
  // the piece bounds of the thread's rows r0 + tid + j WG (rows past r1
  // read r1's), loaded ahead of the partials
  int rs[RT], re[RT];
#pragma unroll
  for (int j = 0; j < RT; ++j) {
    const int r = r0 + tid + j * WG;
    rs[j] = rowptr[min(r, r1)];
    re[j] = rowptr[min(r + 1, r1)];
  }
  if (cnt > CP) {
    Ac a;
    for (int64_t p = wave; p < S; p += NW) {
      const int e1 = o1[p];
      for (int i = o0[p] + lane; i < e1; i += 64) a.add(part[i]);
    }
    xc_finish<A, COMP, WG>(a, wsum, lane, wave);
    __syncthreads();
    if (tid == 0) {
      Ac t;
      t.s = wsum[0];
      t.c = wsum[NW];
      for (int w = 1; w < NW; ++w) {
        Ac o;
        o.s = wsum[w];
        o.c = wsum[NW + w];
        t.merge(o);
      }
      y[r0] = alpha * t.value() + scaled_y(beta, y, r0);
    }
    return;
  }
  if (!scan) {  // too many slices to scan in LDS: a wavefront per run
    for (int64_t p = wave; p < S; p += NW) {
      const int e0 = o0[p], e1 = o1[p];
      for (int i = e0 + lane; i < e1; i += 64) prod[lpos[i]] = part[i];
    }
  } else {
    for (int i = tid; i < CP / 32; i += WG) bits[i] = 0u;
    // scan across the workgroup of (run length | non-empty << 16)
    const int l0 = tid < S ? a1 - a0 : 0;
    const int own = l0 | (l0 > 0 ? 1 << 16 : 0);
    int wsum;
    const int wex = xl_lane_prefix(own, lane, wsum);
    if (lane == 0) wtot[wave] = wsum;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < wave; ++w) wbase += wtot[w];
    // a non-empty run: its first partial index q0 and its rank c among the
    // non-empty runs; its start marked, the words it begins recorded
    const int ex = wbase + wex;
    const int q0 = ex & 0xFFFF, c = ex >> 16;
    if (l0 > 0) {
      delta[c] = a0 - q0;
      atomicOr(&bits[q0 >> 5], 1u << (q0 & 31));
      for (int w = (q0 + 63) >> 6; w * 64 < q0 + l0; ++w) rw[w] = c;
    }
    __syncthreads();
    // partial q = qc + lane, qc = wave 64 + j WG (a whole 64-partial word):
    // its run is the word's first run plus the run starts up to q
    A v[PT];
    int at[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const int qc = wave * 64 + j * WG;  // uniform
      v[j] = A(0);
      at[j] = -1;
      if (qc < cnt) {
        const int w = qc >> 6;
        const uint64_t m = (((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]) & ~1ull;
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const int run = rw[w] + below + (int)((m >> lane) & 1ull);
        const int q = qc + lane;
        if (q < cnt) {
          const int slot = q + delta[run];
          v[j] = part[slot];
          at[j] = lpos[slot];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < PT; ++j)
      if (at[j] >= 0) prod[at[j]] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RT; ++j) {
    const int r = r0 + tid + j * WG;
    const int k0 = rs[j] - base, k1 = re[j] - base;
    const bool is_long = r < r1 && k1 - k0 > 32;
    if (r < r1 && !is_long) {
      Ac s;
      for (int k = k0; k < k1; ++k) s.add(prod[k]);
      y[r] = alpha * s.value() + scaled_y(beta, y, r);
    }
    uint64_t lm = __ballot(is_long);
    while (lm) {
      const int l = __ffsll((unsigned long long)lm) - 1;
      lm &= lm - 1;
      const int rlo = __shfl(k0, l, 64), rhi = __shfl(k1, l, 64);
      Ac a;
      for (int k = rlo + lane; k < rhi; k += 64) a.add(prod[k]);
      a = group_reduce(a, 64);
      if (lane == 0) {
        const int rr = r0 + (tid - lane) + j * WG + l;
        y[rr] = alpha * a.value() + scaled_y(beta, y, rr);
      }
    }
  }
}

// SpMM: Y = alpha*A*X + beta*Y, X n x k, Y m x k (row-major).  One
// wavefront per row; lanes cover the k right-hand sides (k <= 64 per pass),
// so X rows are read as contiguous k-vectors.
template <class VT>
__global__ __launch_bounds__(kWG) void k_csr_spmm(
    const int32_t* __restrict__ rst, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const typename Val<VT>::io* __restrict__ X, int64_t ldx,
    typename Val<VT>::io* __restrict__ Y, int64_t ldy, int64_t m, int k,
    typename Val<VT>::acc alpha, typename Val<VT>::acc beta) {
  using A = typename Val<VT>::acc;
  const int64_t row = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  const int e0 = rst[row], e1 = rst[row + 1];
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int c = c0 + lane;
    if (c >= k) break;
    A s = 0;
    for (int e = e0; e < e1; ++e) s += Val<VT>::get(val, e) * (A)X[(int64_t)col[e] * ldx + c];
    A* yp = &Y[row * ldy + c];
    *yp = alpha * s + scaled_y(beta, yp, 0);
  }
}

template <class VT>
static int run_csr(const DevCSR& d, int64_t m, double alpha, const void* x, double beta,
                   void* y, hipStream_t s, void* scratch) {
  const bool xs_nt = nt_enabled();
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  const bool comp = sizeof(VT) == 2;
  if (d.xl.S > 0) {
    const DevXLds& xl = d.xl;
    AC* part = (AC*)(scratch ? scratch : xl.part);
    if (m == 0) return RVVLA_OK;
    if (xl.nunits > 0) {
      if (comp)
        hipLaunchKernelGGL((k_xl_slices<VT, true>), dim3((unsigned)xl.nwg), dim3(XlShape<VT>::WG), 0,
                           s, xl.col, (const VT*)xl.val, (const int4*)xl.task, xl.units, xl.wg,
                           (const IO*)x, xl.n, part);
      else
        hipLaunchKernelGGL((k_xl_slices<VT, false>), dim3((unsigned)xl.nwg), dim3(XlShape<VT>::WG),
                           0, s, xl.col, (const VT*)xl.val, (const int4*)xl.task, xl.units, xl.wg,
                           (const IO*)x, xl.n, part);
      if (int rc = hip_check(hipGetLastError(), "k_xl_slices launch")) return rc;
    }
    const int nb = (int)xl.nblocks;
    const unsigned grid = (unsigned)(8 * ((nb + 7) / 8));
    auto comb = [&](auto wg) {
      constexpr int WG = decltype(wg)::value;
      if (comp)
        hipLaunchKernelGGL((k_xl_combine<AC, IO, true, WG>), dim3(grid), dim3(WG), 0, s, xl.rowptr,
                           xl.cblk, xl.off, xl.lpos, part, (IO*)y, xl.S, nb, (AC)alpha, (AC)beta);
      else
        hipLaunchKernelGGL((k_xl_combine<AC, IO, false, WG>), dim3(grid), dim3(WG), 0, s,
                           xl.rowptr, xl.cblk, xl.off, xl.lpos, part, (IO*)y, xl.S, nb, (AC)alpha,
                           (AC)beta);
    };
    if (xl.cwg == 256) comb(std::integral_constant<int, 256>());
    else if (xl.cwg == 512) comb(std::integral_constant<int, 512>());
    else comb(std::integral_constant<int, 1024>());
    return hip_check(hipGetLastError(), "k_xl_combine launch");
  }
  if (d.xs.P > 0) {
    const DevXSlice& xs = d.xs;
    IO* xpart = (IO*)(scratch ? scratch : xs.part);
    if (m == 0) return RVVLA_OK;
    if (xs.nblk > 0) {
      if (comp)
        hipLaunchKernelGGL((k_csr_adaptive<VT, true, true>), dim3((unsigned)xs.nblk), dim3(kWG), 0,
                           s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           xpart, (AC)1, (AC)0);
      else if (xs_nt)
        hipLaunchKernelGGL((k_csr_adaptive<VT, false, true>), dim3((unsigned)xs.nblk), dim3(kWG),
                           0, s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           xpart, (AC)1, (AC)0);
      else
        hipLaunchKernelGGL((k_csr_adaptive<VT, false, false>), dim3((unsigned)xs.nblk), dim3(kWG),
                           0, s, xs.rst, xs.col, (const VT*)xs.val, xs.blk, (const IO*)x,
                           xpart, (AC)1, (AC)0);
      if (int rc = hip_check(hipGetLastError(), "k_csr_adaptive (sliced) launch")) return rc;
    }
    const int64_t per_wg = (int64_t)(kWG / 64) * kCombineGroups * 64;
    const unsigned g = (unsigned)((m + per_wg - 1) / per_wg);
    if (comp)
      hipLaunchKernelGGL((k_xs_combine<IO, AC, true>), dim3(g), dim3(kWG), 0, s, xs.mask, xs.base,
                         xs.pref, xpart, (IO*)y, m, xs.nq, xs.P, (AC)alpha, (AC)beta);
    else
      hipLaunchKernelGGL((k_xs_combine<IO, AC, false>), dim3(g), dim3(kWG), 0, s, xs.mask, xs.base,
                         xs.pref, xpart, (IO*)y, m, xs.nq, xs.P, (AC)alpha, (AC)beta);
    return hip_check(hipGetLastError(), "k_xs_combine launch");
  }
  if (d.nblocks == 0) return RVVLA_OK;
  if (comp)
    hipLaunchKernelGGL((k_csr_adaptive<VT, true, false>), dim3((unsigned)d.nblocks), dim3(kWG), 0,
                       s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  else if (xs_nt && sizeof(VT) == 8)  // FP64 entries streamed past L2
    hipLaunchKernelGGL((k_csr_adaptive<VT, false, true>), dim3((unsigned)d.nblocks), dim3(kWG),
                       0, s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  else
    hipLaunchKernelGGL((k_csr_adaptive<VT, false, false>), dim3((unsigned)d.nblocks), dim3(kWG),
                       0, s, d.rst, d.col, (const VT*)d.val, d.blocks, (const IO*)x, (IO*)y,
                       (AC)alpha, (AC)beta);
  return hip_check(hipGetLastError(), "k_csr_adaptive launch");
}

int launch_csr_image(const DevCSR& d, int64_t m, int dtype, double alpha, const void* x,
                     double beta, void* y, hipStream_t s, void* scratch) {
  switch (dtype) {
    case RVVLA_F64: return run_csr<double>(d, m, alpha, x, beta, y, s, scratch);
    case RVVLA_F32: return run_csr<float>(d, m, alpha, x, beta, y, s, scratch);
    case RVVLA_BF16: return run_csr<bf16_t>(d, m, alpha, x, beta, y, s, scratch);
    case RVVLA_F16: return run_csr<_Float16>(d, m, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

int launch_csr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch) {
  return launch_csr_image(A.csr, A.m, A.dtype, alpha, x, beta, y, s, scratch);
}

template <class VT>
static int run_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s) {
  using IO = typename Val<VT>::io;
  using AC = typename Val<VT>::acc;
  if (A.m == 0 || k == 0) return RVVLA_OK;
  const unsigned grid = (unsigned)((A.m + kWG / 64 - 1) / (kWG / 64));
  hipLaunchKernelGGL((k_csr_spmm<VT>), dim3(grid), dim3(kWG), 0, s, A.csr.rst, A.csr.col,
                     (const VT*)A.csr.val, (const IO*)X, ldx, (IO*)Y, ldy, A.m, k, (AC)alpha,
                     (AC)beta);
  return hip_check(hipGetLastError(), "k_csr_spmm launch");
}

int launch_csr_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s) {
  switch (A.dtype) {
    case RVVLA_F64: return run_spmm<double>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_F32: return run_spmm<float>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_BF16: return run_spmm<bf16_t>(A, k, alpha, X, ldx, beta, Y, ldy, s);
    case RVVLA_F16: return run_spmm<_Float16>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  }
  return RVVLA_ERR_VALUE_TYPE;
}

}  // namespace rvvla
