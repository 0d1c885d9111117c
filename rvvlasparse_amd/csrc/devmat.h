// devmat.h -- device-resident matrix images and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "csr_blocks.h"
#include "rvvla_hip.h"

namespace rvvla {

// Column-sliced image of a CSR matrix (see csr_xslice.cpp): P slices (one
// or two per XCD); compressed (row, slice) rows, workgroup blocks as (first, end)
// pairs with block b in slice b mod P, presence masks, per-slice bases and
// per-64-row prefixes, and the partial-sum scratch (one per pair).
struct DevXSlice {
  int P = 0;
  int64_t nblk = 0, nq = 0, pairs = 0;
  const int32_t* rst = nullptr;
  const int32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* blk = nullptr;
  const uint16_t* mask = nullptr;
  const int32_t* base = nullptr;
  const int32_t* pref = nullptr;
  void* part = nullptr;
};

// LDS-sliced image (csr_xlds.cpp): S slices of W columns; the slice pass
// (one workgroup per CU, its LDS holding one slice of x) runs
// tasks of kXlTask entries, one partial per piece; the combine adds each
// row's partials in slice order.
// cwg: threads of a combine workgroup (256, 512 or 1024; combine blocks
// sized for it, see csr_blocks.h).
struct DevXLds {
  int64_t S = 0, W = 0, npieces = 0, ntasks = 0, nunits = 0, n = 0, nwg = 0, nblocks = 0, nent = 0;
  int cwg = kXlCombWG;
  const uint16_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* task = nullptr;
  const int32_t* units = nullptr;
  const int32_t* wg = nullptr;
  const int32_t* rowptr = nullptr;
  const int32_t* cblk = nullptr;
  const int32_t* off = nullptr;
  const uint16_t* lpos = nullptr;
  void* part = nullptr;
};

// CSR: adaptive row blocks.  blocks[b] .. blocks[b+1] is the row range of
// workgroup b (see csr_blocks.h).  When xs.P > 0 the single-vector product
// runs on the column-sliced image; the plain arrays serve SpMM and the
// autograd products.
struct DevCSR {
  const int32_t* rst = nullptr;
  const int32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* blocks = nullptr;
  int64_t nblocks = 0;
  DevXSlice xs;
  DevXLds xl;
  bool sell = false;  // products run on the matrix's SELL-64 image (rvvla_mat_s::sell)
};

// HCSR: one wavefront per unit of work of a 64-row slice of a block-row
// (hcsr_image.cpp): dense runs of 16-bit in-block columns laid 64 lanes
// wide, and a row-ordered tail of (column, lane, value) entries.
struct DevHCSR {
  int64_t R = 0, C = 0, b_m = 0, b_n = 0;
  int64_t nunits = 0, ndrows = 0, ntail = 0, nsplit = 0, nparts = 0;
  const int32_t* unit = nullptr;   // kHcUnitInts per unit
  const int32_t* dxoff = nullptr;  // first column of the block of each slot row
  const uint16_t* dcol = nullptr;  // 64 * ndrows in-block columns (0xFFFF: empty slot)
  const void* dval = nullptr;
  const int32_t* tcol = nullptr;   // ntail tail columns
  const uint8_t* trow = nullptr;   // their lanes (rows inside the slice)
  const void* tval = nullptr;
  const int32_t* split = nullptr;  // per slice cut into units: row0, nrows, first, end part
  void* part = nullptr;            // 64 * nparts partial sums
  // built on the device (rvvla_hcsr_from_device): the reference layout,
  // kept for reading back
  const int32_t* ref_block_ptr = nullptr;
  const uint32_t* ref_row_offset = nullptr;
  const uint32_t* ref_row_ptr = nullptr;
  const uint16_t* ref_col_idx = nullptr;
  const void* ref_values = nullptr;
  int64_t nref_row_ptr = 0;
};

// SELL-C-sigma with C = 64: one wavefront per slice.
struct DevSELL {
  int64_t nslices = 0;
  const int32_t* cl = nullptr;
  const int32_t* cs = nullptr;
  const int32_t* bs = nullptr;
  const int32_t* sbs = nullptr;
  const uint32_t* col = nullptr;
  const uint32_t* perm = nullptr;
  const void* val = nullptr;
  bool perm_identity = false;  // vPerm[s] == s: rows written in place
  // slices wider than kSellPiece steps run as pieces of at most kSellPiece
  // steps, one wavefront each, whose lane partials are added in a fixed order
  int64_t npieces = 0, nsplit = 0;
  const int32_t* pc_slice = nullptr;
  const int32_t* pc_j0 = nullptr;   // step range of each piece
  const int32_t* pc_j1 = nullptr;
  const int32_t* sp_slice = nullptr;  // the split slices
  const int32_t* sp_first = nullptr;  // nsplit + 1: their pieces
  void* part = nullptr;               // npieces * 64 partial sums
};
constexpr int kSellPiece = 64;
// The SELL image of a CSR matrix marks its padding slots (past a row's own
// entries) by this bit of the column; the SELL format keeps plain columns.
constexpr uint32_t kSellPadBit = 0x80000000u;

// CSR5 (omega = 64): tiles as the host container holds them, plus the
// empty-row list and per-tile scratch (leading / trailing partial sums and
// the trailing segment's row).
struct DevCSR5 {
  int64_t ntiles = 0, nempty = 0, tail_base = 0;
  int64_t np = 0, noffsets = 0;  // tile_ptr / empty-offset pointer: np + 1; offsets
  int sigma = 0, dl = 0, bit_y = 0, bit_seg = 0, tail_start = 0, first_tail_row = 0;
  const uint32_t* tile_ptr = nullptr;
  const uint32_t* desc = nullptr;
  const uint32_t* eptr = nullptr;
  const uint32_t* eoff = nullptr;
  const int32_t* rst = nullptr;
  const uint32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* empty_rows = nullptr;
  void* scratch = nullptr;
  // chain[i]: the segment left open at the end of tile i continues through
  // tiles i+1 .. (chain[i] & 0x7FFFFFFF) - 1; bit 31 set: on into the tail.
  const int32_t* chain = nullptr;
  // open segments with more than kCsr5Serial tiles to add, or that run into
  // the tail (tile index t >= 0), and tail rows of more than kCsr5Serial
  // entries (-(row + 1)): finished by a wavefront each instead of a thread
  const int32_t* long_fix = nullptr;
  int64_t nlong = 0;
};
constexpr int kCsr5Serial = 8;

// CVR (SIMD_LEN = 64): one stream per wavefront; records re-laid as a
// flush mask per pack plus the flushed rows in record order.
struct DevCVR {
  int64_t nstreams = 0, nempty = 0;
  int64_t npacks = 0, nrec = 0;         // packs; records in rec_rows
  const int64_t* pack_start = nullptr;  // nstreams + 1
  const int32_t* rec_start = nullptr;   // nstreams
  const uint64_t* flush = nullptr;      // one mask per pack
  // flush targets: a row >= 0 written directly, or -(slot + 2) for a
  // flush of a row recorded more than once (a cut long row): its partial
  // sum goes to part[slot]
  const int32_t* rec_rows = nullptr;
  const int32_t* tail_rows = nullptr;   // 64 per stream, -1 = none
  const uint32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* empty_rows = nullptr;
  // rows recorded more than once: split_rows[i] sums part[split_ptr[2i] ..
  // split_ptr[2i+1]) in slot order (its flushes in record order; slots are
  // given row by row); the first nsplit_long of them have more than 64
  // slots
  int64_t nsplit = 0, nsplit_long = 0, nslots = 0;
  const int32_t* split_rows = nullptr;
  const int32_t* split_ptr = nullptr;
  void* part = nullptr;
};
// CVR flush target of slot q (rows recorded more than once).
constexpr int32_t cvr_slot_target(int64_t q) { return (int32_t)(-q - 2); }

// VHCC (vec_size = 64): one partition per wavefront; the row-end mask of
// every vector, the target rows, and the partition hand-over description.
struct DevVHCC {
  int64_t nparts = 0, nempty = 0;
  int64_t nvec = 0, nends = 0;         // vectors; row ends (row_arr length)
  const int32_t* start_vec = nullptr;  // nparts + 1
  const int32_t* rbase = nullptr;
  const uint64_t* eor = nullptr;       // one row-end mask per vector
  const uint32_t* row_arr = nullptr;   // >= m: the continued row's partial
  const int32_t* fwd_row = nullptr;
  const uint8_t* pass = nullptr;
  const uint32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* empty_rows = nullptr;
  void* scratch = nullptr;             // first_part[nparts], fwd_part[nparts]
};

// LAV (SIMD_Lanes = 64): the dense part's chunk tables (the sparse part
// lives in rvvla_mat_s::csr).
// Chunks with more than kLavPiece steps over all segments are cut into
// pieces of at most kLavPiece steps (within one segment), each run by its
// own wavefront into a 64-lane partial; the chunk's partials are then added
// in an order fixed by the piece count.
constexpr int kLavPiece = 64;
struct DevLAV {
  int64_t nseg = 0, nchunks = 0;
  const int32_t* coff = nullptr;       // nseg * (nchunks + 1) step offsets
  const int64_t* sptr = nullptr;       // nseg segment bases (in steps)
  const uint32_t* out_order = nullptr; // nchunks * 64 rows (shared by all segments)
  const uint32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* cfull = nullptr;      // nseg * nchunks: first step of each chunk that is not full
  const uint16_t* mask = nullptr;      // valid lanes of each step (64 when full)
  int64_t nwhole = 0;                  // chunks run whole
  const int32_t* whole = nullptr;      // their ids (nullptr: all chunks, in order)
  int64_t npieces = 0, nsplit = 0;
  const int32_t* pc_chunk = nullptr;   // chunk of each piece
  const int64_t* pc_k0 = nullptr;      // global step range of each piece
  const int64_t* pc_k1 = nullptr;
  const int32_t* sp_chunk = nullptr;   // the split chunks
  const int32_t* sp_first = nullptr;   // nsplit + 1: their pieces
  void* part = nullptr;                // npieces * 64 partial sums
  // sparse_part rows of at most kLavFuseRow entries: each row's sparse
  // entries are added by the lane that finishes its dense sum (one pass,
  // y written once) instead of by a CSR multiply before the dense kernels
  bool fused = false;
  // built on the device (rvvla_lav_from_device): the rest of the layout,
  // for read-back (chunk_offsets_full_mask, mask, all S copies of
  // out_order) and the dense step count
  bool device_built = false;
  const int32_t* ref_cfull = nullptr;
  const uint16_t* ref_mask = nullptr;
  int64_t total_steps = 0;
};
constexpr int kLavFuseRow = 32;

// VNEC: merge-path partitions of at most vnec_items(accumulator bytes)
// rows + entries, each with its compact list of touched columns (staged in
// LDS).  A workgroup stages at most that many x values and products: 16 KiB
// of LDS, so that ten workgroups fit a CU (1024 items for FP64, 2048 for the
// FP32-accumulating forms).
constexpr int vnec_items(int acc_bytes) { return 8192 / acc_bytes; }
struct DevVNEC {
  int64_t nparts = 0;
  const int32_t* cx = nullptr;       // nparts + 1 merge-path row coordinates
  const int32_t* cy = nullptr;       // nparts + 1 merge-path entry coordinates
  const int32_t* nec_ptr = nullptr;  // nparts + 1
  const int32_t* use_x = nullptr;
  int64_t nlist = 0;                 // entries of use_x
  const uint16_t* ecr = nullptr;     // compacted column rank per entry
  const int32_t* rst = nullptr;
  const void* val = nullptr;
  void* scratch = nullptr;           // first_part[nparts], fwd_part[nparts]
};

// CSC view of a CSR matrix for transposed products: column pointers, row
// of each entry in column order, and the CSR position of that entry.
struct DevCSC {
  const int32_t* ptr = nullptr;   // n + 1
  const int32_t* row = nullptr;   // nnz
  const int32_t* perm = nullptr;  // nnz
};

// Batched persistent launch: per-matrix metadata and the work items.
struct BatchMat {
  int64_t rst_off = 0;   // into the concatenated row pointers (CSR item kernel)
  int64_t ent_off = 0;   // into the concatenated CSR entries
  int64_t sl_off = 0;    // into the concatenated slice tables (SELL item kernel)
  int64_t slot_off = 0;  // into the concatenated SELL slots
  int64_t row_off = 0;   // into the concatenated SELL permutations
  int64_t x_off = 0, y_off = 0;  // multiples of 4 elements
  int32_t m = 0, n = 0, k = 0, fmt = 0;
  int32_t vw = 1;   // right-hand sides per lane: a power of two (<= 4) dividing ld
  int32_t lpe = 1;  // lanes per entry: power of two >= ld / vw
  int32_t gw = 1;   // lanes per row in the CSR item kernel: lpe * slots, <= 64
  int32_t ld = 1;   // row stride of X_i and Y_i: k rounded up to a multiple of vw
};
struct BatchItem {
  int32_t mat, first, count, pad;
};
// One work queue per XCD: queue q holds items[qbase[q] .. qbase[q + 1]);
// its counter is counter[q * kQueueStride] (one cache line each).
constexpr int kMaxQueues = 8;
constexpr int kQueueStride = 32;
struct BatchArgs {
  const BatchMat* mats;
  const BatchItem* items;
  int64_t nitems;
  int* counter;
  int nqueues;
  int64_t qbase[kMaxQueues + 1];
  const int32_t* rst;
  const int32_t* col;
  const int32_t* cl;
  const int32_t* cs;
  const int32_t* scol;
  const int32_t* perm;
};
int launch_batch(int dtype, const BatchArgs& a, const void* val, const void* sval, const void* X,
                 void* Y, double alpha, double beta, int grid, hipStream_t s);

}  // namespace rvvla

struct rvvla_mat_s {
  int format = -1;
  int dtype = -1;
  int64_t m = 0, n = 0, nnz = 0;
  int device = 0;
  bool host_only = false;  // uploaded in host-validation mode (no device copy)
  int64_t device_bytes = 0;
  std::vector<void*> allocs;
  rvvla::DevCSR csr;
  rvvla::DevHCSR hcsr;
  rvvla::DevSELL sell;
  rvvla::DevCSR5 csr5;
  rvvla::DevCVR cvr;
  rvvla::DevVHCC vhcc;
  rvvla::DevLAV lav;
  rvvla::DevVNEC vnec;
  rvvla::DevCSC csc;
  rvvla::DevCSR tcsr;  // CSR of the transpose (built on the first transposed product)
  bool has_tcsr = false;
  // formats other than CSR: builds tcsr from the CSR the matrix was
  // converted from (set by the conversions that have it)
  std::function<void(rvvla_mat_s&)> transpose_source;
  // staging for rvvla_spmv_host
  void* xstage = nullptr;
  void* ystage = nullptr;

  std::vector<void*> temps;  // device buffers of a device-side build

  ~rvvla_mat_s();
  // Allocate `bytes` of device memory, copy `host` into it (if non-null)
  // and keep it for the lifetime of the matrix.
  void* put(const void* host, size_t bytes);
  // Device buffer released by free_temps() (or with the matrix).
  void* temp(size_t bytes);
  void free_temps();
  template <class T>
  const T* put_vec(const std::vector<T>& v) {
    return (const T*)put(v.data(), v.size() * sizeof(T));
  }
};

namespace rvvla {

// Status of a HIP call as an RVVLA code (records the error text).
int hip_check(hipError_t e, const char* what);

int launch_csr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch = nullptr);
// Device-side CSR checks (RVVLA_ERR_INVALID_ARG on a bad CSR) and the
// CSR -> SELL-64 build into M.sell (k_convert.hip); sigma <= 0 picks the
// default rule on the device.
int check_csr_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                     const int32_t* col, hipStream_t s);
template <class V>
int build_sell_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int sigma, hipStream_t s);
// CSR -> LAV (SIMD_Lanes 64) on the device (k_convert_lav.hip): the host
// builder's layout in device memory owned by M; the caller attaches it.
struct LavDeviceHost {
  int64_t S = 0, cc = 0, total = 0, nsp = 0;
  int32_t* srst = nullptr;
  int32_t* scol = nullptr;
  void* sval = nullptr;
  uint32_t* out_order = nullptr;
  int32_t* coff = nullptr;
  int32_t* cfull = nullptr;
  int64_t* sptr = nullptr;
  uint32_t* col_id = nullptr;
  void* dval = nullptr;
  uint16_t* mask = nullptr;
};
template <class V>
int build_lav_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                     const int32_t* col, const V* val, double T, int64_t seg_len,
                     LavDeviceHost& host, hipStream_t s);
// CSR -> VNEC on the device into M.vnec (k_convert_vnec.hip): the host
// builder's partitions, touched-column lists and ranks.
template <class V>
int build_vnec_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t nparts, hipStream_t s);
// CSR -> CVR (SIMD_LEN = 64) on the device into M.cvr (k_convert_cvr.hip):
// the host builder's streams, packs and records.
template <class V>
int build_cvr_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                     const int32_t* col, const V* val, int64_t nstreams, hipStream_t s);
// CSR -> VHCC (vec_size = 64) on the device into M.vhcc
// (k_convert_vhcc.hip): the host builder's partitions, row ends and
// hand-over flags.
template <class V>
int build_vhcc_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t nparts, hipStream_t s);
// CSR -> CSR5 (omega = 64) on the device into M.csr5 (k_convert_csr5.hip):
// the host builder's layout bit for bit, plus the kernels' extras.
template <class V>
int build_csr5_device(rvvla_mat_s& M, int64_t m, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int sigma, hipStream_t s);
// CSR -> HCSR on the device into M.hcsr (k_convert_hcsr.hip): the
// reference layout (bit-identical to the host converter) and the kernels'
// image built from it.
template <class V>
int build_hcsr_device(rvvla_mat_s& M, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                      const int32_t* col, const V* val, int64_t R, int64_t C, double p,
                      hipStream_t s);

// Pieces of the SELL slices wider than kSellPiece steps, from the widths.
void attach_sell_pieces(rvvla_mat_s& M, const std::vector<int>& cl);

// The adaptive CSR product on any CSR image of m rows.  In every product
// launcher, `scratch` (optional) is device memory used in place of the
// image's own partial-sum or hand-over buffer (spmv_scratch_bytes of the
// matrix at least); every word of it that a product reads, the same product
// has written first.
int launch_csr_image(const DevCSR& d, int64_t m, int dtype, double alpha, const void* x,
                     double beta, void* y, hipStream_t s, void* scratch = nullptr);
// Bytes of partial-sum / hand-over buffer a product of A uses (0: none).
int64_t spmv_scratch_bytes(const rvvla_mat_s& A);
// y = alpha A x + beta y in A's format, with `scratch` (or A's own buffers
// when null); spmv_check_args validates the arguments of an SpMV call.
int spmv_dispatch(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                  hipStream_t s, void* scratch);
int spmv_check_args(rvvla_mat_s* A, const void* x, void* y);
int launch_csr_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s);
int launch_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch = nullptr);
int launch_sell(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch = nullptr);
int launch_csr5(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch = nullptr);
int launch_cvr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch = nullptr);
int launch_vhcc(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch = nullptr);
int launch_lav(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s, void* scratch = nullptr);
int launch_vnec(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s, void* scratch = nullptr);
int launch_spmm_backward(const rvvla_mat_s& A, const void* vals, int k, const void* DY,
                         int64_t lddy, const void* X, int64_t ldx, void* dX, int64_t lddx,
                         void* dvals, hipStream_t s);
int launch_spmm_values(const rvvla_mat_s& A, bool transpose, const void* vals, int k, double alpha,
                       const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
                       hipStream_t s);
int launch_sddmm(const rvvla_mat_s& A, int k, const void* DY, int64_t lddy, const void* X,
                 int64_t ldx, double alpha, void* out, hipStream_t s);

}  // namespace rvvla
