/*
 * rvvla_hip.h -- C-ABI of the MI355X (gfx950) SpMV library librvvla_hip.so.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary; every
 * entry point returns an int status (RVVLA_OK == 0, negative on error) that
 * rvvla_strerror() turns into text.
 *
 * The library replaces the reference's multiply path
 *   template<T, Mtx, simd, stage> sparse_mv(op, alpha, const Mtx<T>&, descr,
 *                                          const vector<T>& b, beta, vector<T>& y)
 *   (reference include/spmv_mtx.h:21-29)
 * and its conversion path
 *   template<T, MtxDist, simd> convert(MtxDist<T>& dist, const spMtxCRS<T>& src,
 *                                      const convertParams&)
 *   (reference include/convert.h:30-31)
 * for the eight formats CSR, HCSR, SELL-C-sigma, CSR5, LAV, VNEC, VHCC, CVR.
 * The C++ templates in sparse_matrix.h are implemented on top of this ABI.
 *
 * Device pointers passed to rvvla_spmv/rvvla_spmm must be allocated on the
 * current HIP device; the call is ordered on `stream` (hipStream_t, NULL for
 * the default stream) and does not synchronise.
 */
#ifndef RVVLA_HIP_H
#define RVVLA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* The CSR5 codes mirror reference include/CSR5.h:29-34 (CSR5_* values).     */
#define RVVLA_OK                    0
#define RVVLA_ERR_UNKNOWN_FORMAT   -1  /* CSR5_UNKNOWN_FORMAT               */
#define RVVLA_ERR_UNSUPPORTED_OMEGA -2 /* CSR5_UNSUPPORTED_OMEGA            */
#define RVVLA_ERR_CONVERSION       -3  /* CSR5_CONVERTION_FAILED            */
#define RVVLA_ERR_UNSUPPORTED_SPMV -4  /* CSR5_UNSUPPORTED_SPMV             */
#define RVVLA_ERR_VALUE_TYPE       -5  /* CSR5_UNSUPPORTED_VALUE_TYPE       */
#define RVVLA_ERR_INVALID_ARG      -10 /* std::invalid_argument (crs_to_hcsr.cpp:38) */
#define RVVLA_ERR_LENGTH           -11 /* std::length_error (crs_to_hcsr.cpp:55,87) */
#define RVVLA_ERR_LOGIC            -12 /* std::logic_error (crs_to_hcsr.cpp:91)   */
#define RVVLA_ERR_HIP              -20 /* a HIP runtime call failed           */
#define RVVLA_ERR_NO_DEVICE        -21 /* no HIP device is visible            */
#define RVVLA_ERR_OUT_OF_MEMORY    -22 /* device or host allocation failed    */
#define RVVLA_ERR_NOT_FOUND        -23 /* unknown layout field / parameter    */

/* ---- value storage types ---------------------------------------------- */
/* F64 / F32: the drop-in precisions (x, y, alpha, beta and accumulation in
 * the same type).  BF16 / F16: values stored in 16 bits, x and y in FP32,
 * accumulation in FP32 (compensated on rows longer than 1024 entries).    */
#define RVVLA_F64   0
#define RVVLA_F32   1
#define RVVLA_BF16  2
#define RVVLA_F16   3

/* ---- formats ------------------------------------------------------------ */
#define RVVLA_FMT_CSR   0   /* spMtxCRS          reference include/CRS.h:10   */
#define RVVLA_FMT_HCSR  1   /* spMtxHCSR         reference include/HCSR.h:39  */
#define RVVLA_FMT_SELL  2   /* spMtxSELL_C_Sigma reference include/SELL_C_Sigma.h:23 */
#define RVVLA_FMT_CSR5  3   /* spMtxCSR5         reference include/CSR5.h:57  */
#define RVVLA_FMT_LAV   4   /* spMtxLAV          reference include/LAV.h:10   */
#define RVVLA_FMT_VNEC  5   /* spMtxVNEC         reference include/VNEC.h:35  */
#define RVVLA_FMT_VHCC  6   /* spMtxVHCC         reference include/VHCC.h:43  */
#define RVVLA_FMT_CVR   7   /* spMtxCVR          reference include/CVR.h:22   */
#define RVVLA_NUM_FORMATS 8

typedef struct rvvla_layout_s* rvvla_layout_t;  /* host-side converted matrix */
typedef struct rvvla_mat_s*    rvvla_mat_t;     /* device-resident matrix     */

/* ---- library --------------------------------------------------------- */
const char* rvvla_strerror(int code);
int         rvvla_version(void);
/* Name of a format code ("csr", "hcsr", "sell", "csr5", "lav", "vnec",
 * "vhcc", "cvr"); NULL for an unknown code. */
const char* rvvla_format_name(int format);

/* Measurement helper (no reference counterpart): reads `bytes` (a multiple
 * of 16, 16-byte aligned device buffer) once with streaming loads on
 * `stream`; `sink` is a device buffer of 16 * CUs uint32 that the kernel
 * writes only for a fold value the caller's data will not produce.  Time it
 * with events on `stream`: the device's streaming-read rate, the ceiling of
 * a read-mostly SpMV. */
int rvvla_read_probe(const void* buf, int64_t bytes, void* sink, void* stream);

/* ---- conversion (host) -----------------------------------------------
 * Replaces convert<T, MtxDist, true>(dist, src, params)
 * (reference include/convert.h:30-31; the only reference converter is
 * crs_to_HCSR, src/scalar/crs_to_hcsr.cpp:30-169).
 * `rst`, `col`, `val` are the spMtxCRS fields Rst/Col/Val
 * (reference include/CRS.h:18-20); `val` holds doubles for RVVLA_F64 and
 * floats for every other dtype (BF16/F16 round from FP32 at conversion).
 * `params` is the convertParams map (reference include/convert.h:21-28)
 * written as "key=value" pairs separated by ',' or ';', e.g. "R=2000,C=2000,p=2";
 * NULL or "" keeps every default. */
int rvvla_convert(int format, int dtype, int64_t m, int64_t n, int64_t nnz,
                  const int32_t* rst, const int32_t* col, const void* val,
                  const char* params, rvvla_layout_t* out);
/* Read-only view of one array of the converted host structure, by the
 * reference field name (e.g. "block_ptr", "row_offset", "vPerm",
 * "_csr5_tile_descriptor").  elem_bytes is 1, 2, 4 or 8. */
int rvvla_layout_array(rvvla_layout_t L, const char* field, const void** data,
                       int64_t* count, int* elem_bytes);
/* Scalar field or parameter of the converted structure (e.g. "R", "b_n",
 * "Sigma", "_p", "_tail_tile_start"), returned as a double. */
int rvvla_layout_scalar(rvvla_layout_t L, const char* field, double* value);
int rvvla_layout_free(rvvla_layout_t L);

/* Autotuner (no reference counterpart: the reference leaves the choice of
 * format and of SELL sigma, CSR5 sigma and HCSR R and C to the caller).
 * Times candidate images of the CSR matrix on the current device -- CSR
 * (its chosen image, the plain kernel), SELL (sigma 1, 64, 4096, default),
 * CSR5 (sigma 4, 8, 16), HCSR (the reference's R = C defaults, 2048^2,
 * 4096^2, 256 x 8192, 1024 x 16384), LAV, VNEC, VHCC and CVR with their
 * defaults -- and writes the fastest one's format code to *format and its
 * convertParams text to `out` (`len` bytes).  params: "formats=a+b+..."
 * (default all eight), "reps=N" (timed products a candidate, default 20),
 * "cache=<path>" (a text file of earlier choices, read first and rewritten).
 * Choices are cached for the process by the row-length histogram (rows
 * in power-of-two length buckets to 1/20, the max/mean bucket), the
 * power-of-two buckets of m and n, the dtype and the format list. */
int rvvla_autotune(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                   const int32_t* col, const void* val, const char* params, int* format,
                   char* out, int len);

/* ---- upload (host -> device, once) ------------------------------------ */
int rvvla_upload(rvvla_layout_t L, rvvla_mat_t* out);
/* Direct upload of spMtxCRS fields (reference include/CRS.h:13-20). */
int rvvla_csr_upload(int dtype, int64_t m, int64_t n, int64_t nnz,
                     const int32_t* Rst, const int32_t* Col, const void* Val,
                     rvvla_mat_t* out);
/* Direct upload of spMtxHCSR fields (reference include/HCSR.h:42-53), e.g.
 * a structure produced by the reference's own crs_to_HCSR. */
int rvvla_hcsr_upload(int dtype, int64_t m, int64_t n, int64_t nz,
                      int64_t R, int64_t C, int64_t b_m, int64_t b_n,
                      const int32_t* block_ptr, const uint32_t* row_offset,
                      int64_t row_ptr_len, const uint32_t* row_ptr,
                      const uint16_t* col_idx, const void* values,
                      rvvla_mat_t* out);
/* Direct upload of spMtxSELL_C_Sigma fields
 * (reference include/SELL_C_Sigma.h:26-38); C must be 64. */
int rvvla_sell_upload(int dtype, int64_t m, int64_t n, int64_t nz,
                      int64_t cnt_b, int C, int Sigma,
                      const int32_t* vCl, const int32_t* vCs, const int32_t* vBs,
                      const int32_t* vSBs, int64_t slots, const uint32_t* vCol,
                      const uint32_t* vPerm, const void* vVal, rvvla_mat_t* out);

/* Host-validation mode (per process, off by default): uploads run every
 * check and re-lay of their format but keep the arrays in host memory and
 * need no device; the resulting matrices cannot multiply
 * (RVVLA_ERR_NO_DEVICE).  For validating layouts where no GPU is present. */
int rvvla_set_host_validation(int on);

/* Device-side conversion (no reference counterpart: the reference converts
 * on the host, about 0.3 s for a 4.2 M-row SELL-C-sigma, thousands of
 * products' worth).  Builds the SELL-C-sigma (C = 64) image straight from a
 * CSR already on the device (d_val: double for F64, float otherwise; BF16 /
 * F16 values are rounded on the device) by radix sort, scan and scatter,
 * ordered on `stream`, synchronously.  The CSR is checked on the device
 * first (RVVLA_ERR_INVALID_ARG when bad).  sigma <= 0 applies the default
 * rule (1 when natural 64-row slices pad by at most 1 %; m when sorting all
 * rows at least halves the padded slots; else 16384).  The arrays equal
 * convert(RVVLA_FMT_SELL) + upload bit for bit (the convert path's host
 * fields can be read back with rvvla_mat_read). */
int rvvla_sell_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                           int sigma, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> LAV (SIMD_Lanes 64; the layout of reference
 * include/LAV.h:10-94 that the reference leaves to the caller to build, for
 * a CSR already on the device).  Column histogram, stable radix sorts of
 * columns and rows, scans and scatters give the layout of
 * convert(RVVLA_FMT_LAV) with the same T and segment_len (T < 0 and
 * segment_len <= 0: the convert defaults 0.2 and 2^20), readable with
 * rvvla_mat_read under the spMtxLAV field names ("chunk_offsets",
 * "chunk_offsets_full_mask", "segment_ptr", "out_order", "mask", "val",
 * "col_id", "sparse_part.Rst" / ".Col" / ".Val"), and the kernels' image. */
int rvvla_lav_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                          const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                          double T, int64_t segment_len, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> VNEC (the compaction of reference include/VNEC.h:
 * 78-227, which the reference computes per call on the host, for a CSR
 * already on the device).  nparts merge-path partitions (<= 0: as many
 * rows + entries per partition as one workgroup stages, the default of
 * convert(RVVLA_FMT_VNEC)); per partition the ascending touched-column list
 * and every entry's rank in it, the same arrays as convert + upload
 * (readable with rvvla_mat_read: "cx", "cy" -- the merge-path coordinates
 * -- "nec_ptr", "use_x_indices", "ecr_indices").  The IRD variant selector
 * is not evaluated (one kernel serves its three variants). */
int rvvla_vnec_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                           int64_t nparts, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> VHCC (vec_size 64; the container of reference
 * include/VHCC.h:43-112 that the reference leaves to the caller to fill,
 * for a CSR already on the device).  nparts partitions of (nearly) equal
 * vector counts (<= 0: the default of convert(RVVLA_FMT_VHCC)); per entry
 * its row by binary search, forced row ends at partition ends, the
 * overflow redirect of a continued row, a scan of the row-end flags and
 * one ballot per vector give the arrays of convert + upload, readable with
 * rvvla_mat_read: "start_vec", "rbase", "eor" (one row-end mask per
 * vector), "row_arr", "fwd_row", "pass", "col", "val", "empty_rows". */
int rvvla_vhcc_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                           int64_t nparts, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> CVR (SIMD_LEN 64; the container of reference
 * include/CVR.h:22-81 that the reference leaves to the caller to fill, for a CSR
 * already on the device).  nstreams streams of contiguous rows with
 * balanced entry counts (<= 0: the default of convert(RVVLA_FMT_CVR)); one
 * wavefront per stream replays the greedy lane packing (a count pass, scans,
 * a write pass), giving the arrays of convert + upload, readable with
 * rvvla_mat_read: "pack_start", "rec_start", "flush" (one mask per pack of
 * the lanes that finish a row before it), "rec_rows", "tail_rows" (the rows
 * flushed after a stream's last pack, -1: none), "col", "val",
 * "empty_rows".  Layouts the host builder refuses (rows far longer than a
 * stream's share) are refused with RVVLA_ERR_UNSUPPORTED_SPMV. */
int rvvla_cvr_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                          const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                          int64_t nstreams, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> CSR5 (omega = 64; replaces spMtxCSR5::inputCSR +
 * asCSR5, declared without a body in reference include/CSR5.h:61-65, for a
 * CSR already on the device).  Binary searches, scans and one wavefront per
 * tile build the tile pointers, descriptors, empty offsets and transposed
 * entries -- bit for bit what convert(RVVLA_FMT_CSR5) gives for the same
 * sigma (readable with rvvla_mat_read under the spMtxCSR5 field names) --
 * and the kernels' extras, on `stream`, synchronously.  sigma <= 0: the
 * default (8).  The CSR is checked on the device first. */
int rvvla_csr5_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                           int sigma, void* stream, rvvla_mat_t* out);

/* Device-side CSR -> HCSR (replaces the host converter crs_to_HCSR,
 * reference src/scalar/crs_to_hcsr.cpp:30-169, behind
 * convert<T, spMtxHCSR>, include/convert.h:30-31, for a CSR already on the
 * device).  Radix sort by block, binary searches and scans build the
 * reference layout -- block_ptr, row_offset, row_ptr, col_idx, values, bit
 * for bit what the host converter gives for the same R, C, p (readable with
 * rvvla_mat_read under those names) -- and the kernels' image from it, on
 * `stream`, synchronously.  R, C <= 0 and p <= 0 take the reference
 * defaults (R = C = 2000 for F64, 4000 otherwise; p = 2).  The CSR is
 * checked on the device first (RVVLA_ERR_INVALID_ARG when bad). */
int rvvla_hcsr_from_device(int dtype, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_rst, const int32_t* d_col, const void* d_val,
                           int64_t R, int64_t C, double p, void* stream, rvvla_mat_t* out);

/* Copy of one array of an uploaded matrix into host memory (device or
 * host-validation matrices): SELL "vCl", "vCs", "vBs", "vSBs", "vCol",
 * "vPerm", "vVal"; CSR "rst", "col", "val"; VNEC "cx", "cy", "nec_ptr",
 * "use_x_indices", "ecr_indices"; CSR5 "_csr5_tile_pointer",
 * "_csr5_tile_descriptor", "_csr5_tile_empty_offset_pointer",
 * "_csr5_tile_empty_offset", "_csr_column_index", "_csr_value"; HCSR built
 * on the device "block_ptr", "row_offset", "row_ptr", "col_idx", "values".
 * dst = NULL asks for *count
 * and *elem_bytes only; capacity is in bytes (RVVLA_ERR_LENGTH if short). */
int rvvla_mat_read(rvvla_mat_t A, const char* field, void* dst, int64_t capacity,
                   int64_t* count, int* elem_bytes);

/* Read-only view of one array of a matrix uploaded in host-validation mode
 * (no reference counterpart; for layout tests).  CSR fields: "rst", "col",
 * "blocks" and, when the column-sliced image was built, "xs.rst", "xs.col",
 * "xs.val", "xs.blk", "xs.mask", "xs.base", "xs.pref" ("xs.val" in the
 * storage type of the dtype).  HCSR fields: "brow_ptr", "nb_col", "nb_ptr",
 * "nb_roff", "rp32", "coo16", "col16", "val", "ebc16", "flat_rows",
 * "u_br", "u_lr0", "csplit" (the non-empty block list, the per-entry
 * arrays, the entry-by-entry block-rows and the row units of the others).
 * *count is 0 for an absent array.  Matrices on a device give
 * RVVLA_ERR_LOGIC. */
int rvvla_mat_view(rvvla_mat_t A, const char* field, const void** data, int64_t* count,
                   int* elem_bytes);

int rvvla_mat_free(rvvla_mat_t A);
/* Shape and storage of a device matrix.  device_bytes = bytes of the
 * format's arrays resident in HBM (what one multiply streams, x/y apart). */
int rvvla_mat_info(rvvla_mat_t A, int64_t* m, int64_t* n, int64_t* nnz,
                   int* format, int* dtype, int64_t* device_bytes);

/* ---- multiply ------------------------------------------------------------
 * y = alpha * A * x + beta * y  (the MULT stage of sparse_mv,
 * reference include/types.h:33-38).  x and y are device pointers of the
 * compute type (double for F64, float for F32/BF16/F16).  With beta == 0
 * y is write-only (BLAS convention); the reference's scalar CSR reads y
 * there too (spmv_mtx_crs.cpp:37-43), so the two differ only when y already
 * holds NaN or Inf.  Products are ordered on `stream`; some formats keep
 * per-matrix partial sums on the device (CSR5, VHCC, VNEC, the column- and
 * LDS-sliced CSR images, the split slices of SELL and chunks of LAV), so products of one
 * matrix must not run concurrently on different streams.  The environment
 * variable RVVLA_NT=0 turns off the nontemporal loads of the FP64 SELL and
 * CSR kernels (for comparisons). */
int rvvla_spmv(rvvla_mat_t A, double alpha, const void* x, double beta,
               void* y, void* stream);
/* rvvla_spmv without any allocation: no reference counterpart; for callers
 * that cannot allocate device memory during a product (code under stream
 * capture, tight serving loops, harnesses that forbid it).  The same kernels
 * and results as rvvla_spmv, but the per-product partial-sum / hand-over
 * buffers live in one 16 MiB device array that the library holds on each
 * device until the process ends (its address is looked up on the first
 * call on a device).  A product needing more than 16 MiB of such buffers
 * returns RVVLA_ERR_LENGTH.  All calls on one device share the array:
 * calls that overlap on one device race. */
int rvvla_spmv_noalloc(rvvla_mat_t A, double alpha, const void* x, double beta,
                       void* y, void* stream);
/* The product with the operation of sparse_mv's type_op
 * (reference include/types.h:24-27; the reference accepts the flag but
 * always computes A x): op 0 = SPARSE_OPERATION_NON_TRANSPOSE (rvvla_spmv),
 * op 1 = SPARSE_OPERATION_TRANSPOSE, y = alpha * A^T * x + beta * y with x of
 * m entries and y of n, every format and dtype: the first transposed
 * product builds a CSR image of A^T on the device (counting sort by column,
 * about as much memory again as A) -- for CSR from the matrix itself, for
 * the other formats from the CSR they were converted from (rvvla_upload of
 * an rvvla_convert layout keeps a reference to it; the C++ convert keeps a
 * host copy).  Matrices uploaded from format arrays (rvvla_<fmt>_upload) or
 * built on the device (rvvla_<fmt>_from_device) have no CSR to build it
 * from and return RVVLA_ERR_UNSUPPORTED_SPMV. */
int rvvla_spmv_op(rvvla_mat_t A, int op, double alpha, const void* x, double beta,
                  void* y, void* stream);
/* Y = alpha * A * X + beta * Y for k right-hand sides; X is n x k and Y is
 * m x k, row-major with leading dimensions ldx >= k and ldy >= k
 * (CSR, F32/F64).  No reference counterpart: the reference takes one vector. */
int rvvla_spmm(rvvla_mat_t A, int k, double alpha, const void* X, int64_t ldx,
               double beta, void* Y, int64_t ldy, void* stream);
/* ---- differentiable SpMM (CSR, F32 / F64) ------------------------------
 * No reference counterpart (the reference multiplies one vector with fixed
 * values); these back the PyTorch autograd extension.
 * rvvla_spmm_values: Y = alpha * A(vals) X + beta * Y with A's CSR
 *   structure and the caller's values `vals` (device, CSR order; NULL = the
 *   uploaded values).
 * rvvla_spmm_transpose: Y = alpha * A(vals)^T X + beta * Y (Y is n x k)
 *   through the cached CSC permutation -- gather-based, no atomics.
 * rvvla_sddmm: out[e] = alpha * <DY[row_e, :], X[col_e, :]> for every entry
 *   e (the gradient of the values of A in Y = A X).
 * rvvla_prepare_transpose builds and uploads the CSC permutation (the first
 * rvvla_spmm_transpose call does it otherwise; it allocates and copies
 * synchronously, so do it before capturing a stream into a graph). */
int rvvla_prepare_transpose(rvvla_mat_t A);
int rvvla_spmm_values(rvvla_mat_t A, const void* vals, int k, double alpha, const void* X,
                      int64_t ldx, double beta, void* Y, int64_t ldy, void* stream);
int rvvla_spmm_transpose(rvvla_mat_t A, const void* vals, int k, double alpha, const void* X,
                         int64_t ldx, double beta, void* Y, int64_t ldy, void* stream);
int rvvla_sddmm(rvvla_mat_t A, int k, double alpha, const void* DY, int64_t lddy, const void* X,
                int64_t ldx, void* out_vals, void* stream);
/* The backward of Y = A(vals) X in one pass over A's columns (no reference
 * counterpart): dX = A(vals)^T DY (n x k, leading dimension lddx, written,
 * not accumulated) and dvals[e] = <DY[row_e, :], X[col_e, :]> (CSR order),
 * each DY row gathered once for both; either output may be NULL.  Builds
 * the cached CSC structure on first use (as rvvla_spmm_transpose).  Same
 * results as rvvla_spmm_transpose (alpha 1, beta 0) plus rvvla_sddmm
 * (alpha 1), up to the order of additions. */
int rvvla_spmm_backward(rvvla_mat_t A, const void* vals, int k, const void* DY, int64_t lddy,
                        const void* X, int64_t ldx, void* dX, int64_t lddx, void* dvals,
                        void* stream);

/* ---- batched uneven SpMM in one persistent launch ----------------------
 * No reference counterpart.  `count` CSR matrices (host arrays; values
 * double for F64, float for F32), matrix i with k[i] in 1..16 right-hand
 * sides, become one device work queue.  The batch owns row-major X_i
 * (n_i x k_i) and Y_i (m_i x k_i) blocks in HBM with row stride ld_i
 * (rvvla_batch_block gives their device addresses and ld_i: k_i rounded up
 * to a multiple of 4 (FP64: 2) when k_i > 2, so that each lane reads 16
 * bytes of X_i at once; columns k_i .. ld_i-1 are padding, zero after
 * creation, never written, and do not affect Y_i).  rvvla_batch_run
 * computes every Y_i = alpha A_i X_i + beta Y_i with one persistent kernel
 * (plus a one-word queue reset); rvvla_batch_capture records that pair into
 * a hipGraph once and rvvla_batch_replay launches it.  params: "tune=0|1" (default 1: the
 * autotuner times the candidate item kernels on the first matrix of each
 * row-length-histogram key), "format=0|1,param=P" to force CSR rows (P rows
 * per item) or SELL-64 (sigma P), "cache=<path>" to persist the tuner's
 * choices, "wave_items=W" (default 128) work items per wavefront of the
 * persistent grid. */
typedef struct rvvla_batch_s* rvvla_batch_t;
int rvvla_batch_create(int dtype, int count, const int64_t* m, const int64_t* n,
                       const int32_t* const* rst, const int32_t* const* col,
                       const void* const* val, const int* k, const char* params,
                       rvvla_batch_t* out);
/* Device addresses of X_i and Y_i, their row stride ld_i (elements), and
 * the item kernel chosen for matrix i (RVVLA_FMT_CSR or RVVLA_FMT_SELL).
 * Any output pointer may be NULL. */
int rvvla_batch_block(rvvla_batch_t B, int i, void** x, void** y, int64_t* ld, int* format);
int rvvla_batch_tune_key(rvvla_batch_t B, int i, char* buf, int len);
int rvvla_batch_run(rvvla_batch_t B, double alpha, double beta, void* stream);
int rvvla_batch_capture(rvvla_batch_t B, double alpha, double beta);
int rvvla_batch_replay(rvvla_batch_t B, void* stream);
int rvvla_batch_free(rvvla_batch_t B);

/* ALL_STAGES on host vectors: upload b (and y), multiply, download y,
 * synchronously -- the std::vector form of sparse_mv
 * (reference include/spmv_mtx.h:21-29). */
int rvvla_spmv_host(rvvla_mat_t A, double alpha, const void* b_host,
                    double beta, void* y_host);
/* The three stages of that call, separately (reference
 * include/types.h:33-38): PREPARATION copies b and y into the matrix's
 * device staging vectors (y_host may be NULL: y then starts at zero),
 * MULT multiplies them in place, GET_RESULTS copies y back.  All three are
 * synchronous; the staging vectors belong to the matrix, so concurrent
 * host-vector calls on one matrix must be serialised by the caller. */
int rvvla_stage_in(rvvla_mat_t A, const void* b_host, const void* y_host);
int rvvla_mult_staged(rvvla_mat_t A, double alpha, double beta);
int rvvla_stage_out(rvvla_mat_t A, void* y_host);
/* The same with an operation (op as in rvvla_spmv_op; for op 1, b has m
 * entries and y has n). */
int rvvla_spmv_host_op(rvvla_mat_t A, int op, double alpha, const void* b_host,
                       double beta, void* y_host);
int rvvla_stage_in_op(rvvla_mat_t A, int op, const void* b_host, const void* y_host);
int rvvla_mult_staged_op(rvvla_mat_t A, int op, double alpha, double beta);
int rvvla_stage_out_op(rvvla_mat_t A, int op, void* y_host);

/* ---- synthetic matrices (host, deterministic, OpenMP) ----------------
 * Generators for the benchmark matrices.  Every array is allocated by the
 * library and released with rvvla_gen_free.  Values are written as double
 * (convert them for other dtypes).  Streams: splitmix64 keyed by
 * (seed, purpose, row), so results do not depend on the thread count. */
typedef struct {
  int64_t m, n, nnz;
  int32_t* rst;
  int32_t* col;
  double*  val;
} rvvla_gen_csr;
/* 5-point Laplacian on a g x g grid, natural row-major order, 4 on the
 * diagonal and -1 off it. */
int rvvla_gen_laplacian5(int64_t g, rvvla_gen_csr* out);
/* n x n, each row = diagonal + (k-1) distinct uniform off-diagonal columns
 * (include_diag=1), or k distinct uniform columns (include_diag=0); sorted
 * columns; values uniform[-1,1] (normal=0) or N(0, stdev^2) (normal=1). */
int rvvla_gen_random_rows(int64_t n, int k, int include_diag, int normal,
                          double stdev, uint64_t seed, rvvla_gen_csr* out);
/* n x n, row lengths uniform in [kmin, kmax], distinct uniform columns,
 * sorted; values N(0,1). */
int rvvla_gen_random_rows_var(int64_t n, int kmin, int kmax, uint64_t seed,
                              rvvla_gen_csr* out);
/* n x n power-law rows: length ~ discrete power law exponent `alpha` on
 * [1, maxlen], rescaled (stochastic rounding, clipped to [1, maxlen]) so the
 * expected total is `target_nnz`; half of each row within +-`band` of the
 * diagonal (capped by the columns available there), the rest uniform;
 * values N(0,1). */
int rvvla_gen_powerlaw(int64_t n, double alpha, int64_t maxlen, int64_t target_nnz,
                       int64_t band, uint64_t seed, rvvla_gen_csr* out);
/* v[i] = uniform[lo, hi) from (seed, purpose, i). */
int rvvla_gen_uniform(int64_t count, double lo, double hi, uint64_t seed,
                      uint64_t purpose, double* v);
int rvvla_gen_free(rvvla_gen_csr* g);

/* ---- matrix files (host) ---------------------------------------------
 * The reference keeps a Matrix Market type code in spMtxCRS
 * (reference include/CRS.h:8,17) and names a mtx-to-bin converter in its
 * build options (reference cmake/TestEnvironment.cmake:62) without shipping
 * either; these are them.  Results are canonical CSR (columns ascending in
 * each row, duplicate entries summed) in arrays released with
 * rvvla_gen_free; matcode (may be NULL) gets the four MM_typecode letters.
 *
 * rvvla_mtx_read: Matrix Market coordinate files, field real / integer /
 * pattern (value 1), symmetry general / symmetric / skew-symmetric (the
 * mirror of each off-diagonal entry is added).  Errors: RVVLA_ERR_NOT_FOUND
 * (cannot open), RVVLA_ERR_INVALID_ARG (malformed), RVVLA_ERR_UNSUPPORTED_SPMV
 * (array format, complex, hermitian), RVVLA_ERR_LENGTH (beyond int32).
 *
 * rvvla_bin_write / rvvla_bin_read: binary CSR, little-endian:
 * "RVVLACSR", int32 version 1, char[4] matcode, int64 m, n, nnz, then
 * int32 Rst[m+1], int32 Col[nnz], float64 Val[nnz].  A file that does not
 * hold a valid CSR is refused with RVVLA_ERR_INVALID_ARG. */
int rvvla_mtx_read(const char* path, rvvla_gen_csr* out, char matcode[4]);
int rvvla_bin_write(const char* path, const rvvla_gen_csr* a, const char matcode[4]);
int rvvla_bin_read(const char* path, rvvla_gen_csr* out, char matcode[4]);

#ifdef __cplusplus
}
#endif
#endif /* RVVLA_HIP_H */
