// capi.cpp -- the extern "C" boundary of librvvla_hip.so.
//
// Conversion builds the reference host containers (include/*.h), upload
// validates every index against the matrix shape on the host (so no kernel
// can be handed an out-of-range gather) and copies the device image once,
// and rvvla_spmv dispatches to the format's gfx950 kernel on the caller's
// stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdlib>
#include <cstdio>
#include <functional>
#include <mutex>

#include "devmat.h"
#include "internal.h"

using namespace SparseMatrixLib;

// ---------------------------------------------------------------- errors
namespace rvvla {

static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
const std::string& last_error() { return g_last_error; }

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return RVVLA_OK;
  set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? RVVLA_ERR_OUT_OF_MEMORY : RVVLA_ERR_HIP;
}

SparseMatrixLib::convertParams parse_params(const char* text) {
  convertParams p;
  if (!text) return p;
  std::string s(text), item;
  auto flush = [&](const std::string& it) {
    std::string t;
    for (char ch : it)
      if (!std::isspace((unsigned char)ch)) t += ch;
    if (t.empty()) return;
    const size_t eq = t.find('=');
    if (eq == std::string::npos || eq == 0)
      throw Error(RVVLA_ERR_INVALID_ARG, "params: expected key=value, got '" + t + "'");
    const std::string key = t.substr(0, eq), v = t.substr(eq + 1);
    char* end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (end == v.c_str() || *end != 0)
      throw Error(RVVLA_ERR_INVALID_ARG, "params: value of '" + key + "' is not a number");
    convertValue cv;
    // integers are stored in .i, everything else in .d (the reference map
    // is a union; each converter knows which member its keys use)
    if (v.find_first_of(".eE") == std::string::npos && d >= -2147483648.0 && d <= 2147483647.0)
      cv.i = (int)d;
    else
      cv.d = d;
    p.param[key] = cv;
    // remember the textual form so double-valued keys read correctly
    convertValue dv;
    dv.d = d;
    p.param["#" + key] = dv;
  };
  for (char ch : s) {
    if (ch == ',' || ch == ';') {
      flush(item);
      item.clear();
    } else {
      item += ch;
    }
  }
  flush(item);
  return p;
}

int param_int(const convertParams& p, const char* key, int dflt) {
  auto it = p.param.find(std::string("#") + key);
  if (it != p.param.end()) return (int)it->second.d;
  auto jt = p.param.find(key);
  return jt == p.param.end() ? dflt : jt->second.i;
}

double param_double(const convertParams& p, const char* key, double dflt) {
  auto it = p.param.find(std::string("#") + key);
  if (it != p.param.end()) return it->second.d;
  auto jt = p.param.find(key);
  return jt == p.param.end() ? dflt : jt->second.d;
}

uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

uint16_t f32_to_f16(float f) {
  const _Float16 h = (_Float16)f;  // round to nearest even
  uint16_t u;
  std::memcpy(&u, &h, 2);
  return u;
}

float bf16_to_f32(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

float f16_to_f32(uint16_t b) {
  _Float16 h;
  std::memcpy(&h, &b, 2);
  return (float)h;
}

}  // namespace rvvla

using rvvla::Error;

// ---------------------------------------------------------------- device matrix
namespace rvvla {
static std::atomic<int> g_host_validation{0};
}

rvvla_mat_s::~rvvla_mat_s() {
  free_temps();
  for (void* p : allocs) {
    if (host_only) std::free(p);
    else (void)hipFree(p);
  }
  if (xstage) (void)hipFree(xstage);
  if (ystage) (void)hipFree(ystage);
}

void* rvvla_mat_s::temp(size_t bytes) {
  void* d = nullptr;
  const int rc = rvvla::hip_check(hipMalloc(&d, bytes ? bytes : 8), "hipMalloc");
  if (rc) throw Error(rc, rvvla::last_error());
  temps.push_back(d);
  return d;
}

void rvvla_mat_s::free_temps() {
  for (void* p : temps) (void)hipFree(p);
  temps.clear();
}

void* rvvla_mat_s::put(const void* host, size_t bytes) {
  void* d = nullptr;
  const size_t nb = bytes ? bytes : 4;
  if (host_only) {  // host-validation mode: same copies, host memory
    d = std::malloc(nb);
    if (!d) throw Error(RVVLA_ERR_OUT_OF_MEMORY, "host allocation");
    allocs.push_back(d);
    device_bytes += (int64_t)bytes;
    if (host && bytes) std::memcpy(d, host, bytes);
    return d;
  }
  int rc = rvvla::hip_check(hipMalloc(&d, nb), "hipMalloc");
  if (rc) throw Error(rc, rvvla::last_error());
  allocs.push_back(d);
  device_bytes += (int64_t)bytes;
  if (host && bytes) {
    rc = rvvla::hip_check(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    if (rc) throw Error(rc, rvvla::last_error());
  }
  return d;
}

namespace rvvla {

static int check_dtype_for(int dtype, bool is_double) {
  if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
  if ((dtype == RVVLA_F64) != is_double)
    throw Error(RVVLA_ERR_VALUE_TYPE, "F64 matrices hold doubles, every other dtype floats");
  return dtype;
}

// Value array in the device storage type of `dtype`.
template <class V>
static const void* put_values(rvvla_mat_s& M, int dtype, const V* v, size_t count) {
  if (dtype == RVVLA_BF16 || dtype == RVVLA_F16) {
    std::vector<uint16_t> h(count);
    for (size_t i = 0; i < count; ++i)
      h[i] = dtype == RVVLA_BF16 ? f32_to_bf16((float)v[i]) : f32_to_f16((float)v[i]);
    return M.put_vec(h);
  }
  return M.put(v, count * sizeof(V));
}

static std::unique_ptr<rvvla_mat_s> new_mat(int format, int dtype, int64_t m, int64_t n,
                                            int64_t nnz) {
  int dev = -1;
  const bool host_only = g_host_validation.load() != 0;
  if (!host_only) {
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) throw Error(RVVLA_ERR_NO_DEVICE, last_error());
  }
  auto M = std::make_unique<rvvla_mat_s>();
  M->host_only = host_only;
  M->format = format;
  M->dtype = dtype;
  M->m = m;
  M->n = n;
  M->nnz = nnz;
  M->device = dev;
  return M;
}

static void check_csr(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col) {
  if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
    throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
  if (m > 0 && !rst) throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null row pointer");
  if (nnz > 0 && !col) throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null column array");
  if (m > 0 && (rst[0] != 0 || rst[m] != nnz))
    throw Error(RVVLA_ERR_INVALID_ARG, "CSR: Rst[0] must be 0 and Rst[m] == nnz");
  for (int64_t i = 0; i < m; ++i)
    if (rst[i + 1] < rst[i]) throw Error(RVVLA_ERR_INVALID_ARG, "CSR: Rst not monotone");
  for (int64_t e = 0; e < nnz; ++e)
    if (col[e] < 0 || col[e] >= n) throw Error(RVVLA_ERR_INVALID_ARG, "CSR: column out of range");
}

void check_csr_public(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col) {
  check_csr(m, n, nnz, rst, col);
}

// XCDs of the current device (8 assumed in host-validation mode).
static int device_xcds(const rvvla_mat_s& M) {
  int xcc = 8;
  if (!M.host_only &&
      hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, M.device) != hipSuccess)
    xcc = 1;
  return std::max(1, std::min(xcc, 8));
}

// Number of column slices for the sliced CSR image: `request` > 0 is taken
// as given (1 = no slicing), otherwise (-1) slicing is used when x is
// larger than one XCD's L2 (4 MiB) and more than a fifth of the entries lie
// over 65536 columns away from the row's diagonal position: one slice per
// XCD, or two per XCD when a slice of x would exceed the 4 MiB L2.
template <class V>
static void fill_sell(rvvla_mat_s& M, int dtype, const spMtxSELL_C_Sigma<V>& a,
                      bool keep_perm = true);

// The SELL-64 image of a CSR matrix (rows in natural order, sigma = 1):
// `request` 1 = on, 0 = off, -1 = chosen from the matrix: at least 65536
// rows, none longer than kSellPiece entries (so every slice runs as one
// wavefront), whose 64-row slices pad by at most 1 % of the entries and
// hold fewer than 2^31 slots (the 5-point Laplacian 2048^2: 53 us against
// 59 us for the adaptive CSR kernel on MI355X, FP64).  The SELL kernel adds
// a row's entries in four interleaved partial sums (entries j mod 4), a
// fixed order within the usual bound, and reads only the row's own entries
// (the row lengths bound each lane), as the CSR product does.
template <class V>
static bool choose_csr_sell(const spMtxCRS<V>& a, int request) {
  if (request >= 0) return request > 0 && a.m > 0;
  const int64_t m = (int64_t)a.m, nnz = (int64_t)a.nz;
  if (m < 65536 || nnz == 0) return false;
  int64_t padded = 0;
  for (int64_t r0 = 0; r0 < m; r0 += 64) {
    const int64_t r1 = std::min<int64_t>(m, r0 + 64);
    int64_t w = 0;
    for (int64_t r = r0; r < r1; ++r) w = std::max<int64_t>(w, a.Rst[r + 1] - a.Rst[r]);
    if (w > kSellPiece) return false;
    padded += w * (r1 - r0);
  }
  if (padded > 2147483647LL) return false;
  return (double)(padded - nnz) <= 0.01 * (double)nnz;
}

template <class V>
static int choose_xslices(const rvvla_mat_s& M, const spMtxCRS<V>& a, int request) {
  if (request > 16) throw Error(RVVLA_ERR_INVALID_ARG, "CSR: xslices must be at most 16");
  if (request == 0 || request == 1) return 0;
  if (request > 1) return request;
  const int xcc = device_xcds(M);
  const int64_t xbytes = (int64_t)a.n * (M.dtype == RVVLA_F64 ? 8 : 4);
  if (xcc < 2 || xbytes <= (int64_t(4) << 20)) return 0;
  if (far_fraction((int64_t)a.m, (int64_t)a.n, a.Rst, a.Col, 65536) <= 0.2) return 0;
  return xbytes / xcc > (int64_t(4) << 20) ? 2 * xcc : xcc;
}

// CUs of the current device (256 assumed in host-validation mode).
static int device_cus(const rvvla_mat_s& M) {
  int cus = 256;
  if (!M.host_only &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, M.device) != hipSuccess)
    cus = 256;
  return std::max(1, cus);
}

// The LDS-sliced image: `request` 1 = on, 0 = off, -1 = chosen from the
// matrix: where the column-sliced image would be chosen by default (an
// explicit xslices keeps the column-sliced image), with at least one slice
// width of entries per workgroup (to amortise loading x into LDS) and at
// least 3 entries per (row, slice) pair (each pair costs a partial written
// and read back).
template <class V>
static bool choose_xlds(const rvvla_mat_s& M, const spMtxCRS<V>& a, int request, int xslices) {
  if (request >= 0) return request > 0 && a.n > 0 && a.m > 0;
  if (xslices >= 0 || a.n == 0 || a.m == 0) return false;
  if (choose_xslices(M, a, -1) == 0) return false;
  const int64_t W = xl_slice_cols(M.dtype == RVVLA_F64 ? 8 : 4);
  if ((int64_t)a.nz < (int64_t)device_cus(M) * W) return false;
  return (int64_t)a.nz >= 3 * xlds_pairs((int64_t)a.m, a.Rst, a.Col, W);
}

template <class V>
rvvla_mat_t upload_csr(int dtype, const spMtxCRS<V>& a, int xslices, int xlds, int sell, int xlcomb) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  check_csr((int64_t)a.m, (int64_t)a.n, (int64_t)a.nz, a.Rst, a.Col);
  auto M = new_mat(RVVLA_FMT_CSR, dtype, (int64_t)a.m, (int64_t)a.n, (int64_t)a.nz);
  const std::vector<int32_t> blocks = csr_row_blocks(a.Rst, 0, (int64_t)a.m);
  M->csr.rst = (const int32_t*)M->put(a.Rst, (a.m + 1) * sizeof(int));
  M->csr.col = (const int32_t*)M->put(a.Col, a.nz * sizeof(int));
  M->csr.val = put_values(*M, dtype, a.Val, a.nz);
  M->csr.blocks = M->put_vec(blocks);
  M->csr.nblocks = (int64_t)blocks.size() / 2;
  if (choose_xlds(*M, a, xlds, xslices)) {
    XLdsHost<V> X;
    const int io = dtype == RVVLA_F64 ? 8 : 4;
    const int64_t W = xl_slice_cols(io);
    // the combine's run table holds (blocks + 1) * slices offsets
    const int cwg = xlcomb > 0 ? xlcomb : kXlCombWG;
    if (cwg != 256 && cwg != 512 && cwg != 1024)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: xlcomb must be 256, 512 or 1024");
    const int64_t crows = xl_comb_rows(cwg), cpieces = xl_comb_pieces(io, cwg);
    const int64_t est_blocks = (int64_t)a.m / crows + 2 * (int64_t)a.nz / cpieces + 2;
    if (est_blocks * (((int64_t)a.n + W - 1) / W) > (int64_t(1) << 27))
      throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "CSR: LDS-sliced image run table too large");
    const int vpw = dtype == RVVLA_F64 ? 2 : dtype == RVVLA_F32 ? 4 : 8;
    // FP64 (one 512-thread workgroup a CU) splits its tasks over four
    // workgroups a CU, so that CUs that finish early take more (measured on
    // the 4.2M-row power-law matrix: 244 -> 239 us); the 32-bit-x forms
    // keep one a CU (more measured slower there)
    const int64_t nwg_x = (int64_t)device_cus(*M) * (dtype == RVVLA_F64 ? 4 : 1);
    build_xlds<V>(X, (int64_t)a.m, (int64_t)a.n, a.Rst, a.Col, a.Val, W, nwg_x, crows,
                  cpieces, vpw);
    DevXLds& d = M->csr.xl;
    d.cwg = cwg;
    d.S = X.S;
    d.W = X.W;
    d.n = (int64_t)a.n;
    d.npieces = X.npieces;
    d.nent = X.nent;
    d.ntasks = (int64_t)X.task.size() / 4;
    d.nunits = (int64_t)X.units.size() / 3;
    d.nwg = (int64_t)X.wg.size() - 1;
    d.nblocks = (int64_t)X.cblk.size() / 2;
    d.col = M->put_vec(X.col);
    d.val = put_values(*M, dtype, X.val.data(), X.val.size());
    d.task = M->put_vec(X.task);
    d.units = M->put_vec(X.units);
    d.wg = M->put_vec(X.wg);
    d.rowptr = M->put_vec(X.rowptr);
    d.cblk = M->put_vec(X.cblk);
    d.off = M->put_vec(X.off);
    d.lpos = M->put_vec(X.lpos);
    d.part = M->put(nullptr, (size_t)std::max<int64_t>(d.npieces, 1) * io);
    return M.release();
  }
  const int P = choose_xslices(*M, a, xslices);
  if (P == 0 && choose_csr_sell(a, sell)) {
    spMtxSELL_C_Sigma<V> S;
    build_sell<V>(S, a, 64, 1);
    // padding slots (past a row's own entries) carry kSellPadBit in their
    // column, so the kernel drops their products without reading Rst
    for (int64_t i = 0; i < (int64_t)S.cnt_b; ++i) {
      const int64_t h = S.vBs[(size_t)i], w = S.vCl[(size_t)i], c0 = S.vCs[(size_t)i];
      const int64_t r0 = S.vSBs[(size_t)i];
      for (int64_t k = 0; k < h; ++k) {
        const int64_t len = a.Rst[r0 + k + 1] - a.Rst[r0 + k];
        for (int64_t j = len; j < w; ++j) S.vCol[(size_t)(c0 + j * h + k)] |= kSellPadBit;
      }
    }
    fill_sell(*M, dtype, S, false);
    M->csr.sell = true;
  }
  if (P > 0) {
    XSliceHost<V> X;
    build_xslice<V>(X, (int64_t)a.m, (int64_t)a.n, a.Rst, a.Col, a.Val, P,
                    std::min(P, device_xcds(*M)));
    DevXSlice& d = M->csr.xs;
    d.P = P;
    d.nq = X.nq;
    d.pairs = X.base[(size_t)P];
    d.nblk = (int64_t)X.blk.size() / 2;
    d.rst = M->put_vec(X.rst);
    d.col = M->put_vec(X.col);
    d.val = put_values(*M, dtype, X.val.data(), X.val.size());
    d.blk = M->put_vec(X.blk);
    d.mask = M->put_vec(X.mask);
    d.base = M->put_vec(X.base);
    d.pref = M->put_vec(X.pref);
    d.part = M->put(nullptr, (size_t)d.pairs * (dtype == RVVLA_F64 ? 8 : 4));
  }
  return M.release();
}

// Uploads the image of hcsr_image.cpp (from the validated layout) into
// M->hcsr; values as V (double for F64, float otherwise) narrowed for
// BF16/F16.
template <class V>
void put_hcsr_image(rvvla_mat_s& M, int dtype, const HcsrImage<V>& im) {
  DevHCSR& d = M.hcsr;
  d.nunits = im.nunits;
  d.ndrows = (int64_t)im.dcol.size() / 64;
  d.ntail = (int64_t)im.tcol.size();
  d.nsplit = im.nsplit;
  d.nparts = im.nparts;
  d.unit = M.put_vec(im.unit);
  d.dxoff = M.put_vec(im.dxoff);
  d.dcol = M.put_vec(im.dcol);
  d.dval = put_values(M, dtype, im.dval.data(), im.dval.size());
  d.tcol = M.put_vec(im.tcol);
  d.trow = M.put_vec(im.trow);
  d.tval = put_values(M, dtype, im.tval.data(), im.tval.size());
  d.split = M.put_vec(im.split);
  const size_t pb = (size_t)im.nparts * 64 * (dtype == RVVLA_F64 ? 8 : 4);
  d.part = pb ? M.put(nullptr, pb) : nullptr;
}

// The image of a layout built on the device (rvvla_hcsr_from_device): the
// layout is copied to the host (16-bit values widened exactly) and laid
// out by build_hcsr_image.
template <class V>
void hcsr_image_from_device(rvvla_mat_s& M, hipStream_t s) {
  DevHCSR& d = M.hcsr;
  const int64_t nb = d.b_m * d.b_n + 1, nz = M.nnz;
  std::vector<int32_t> bp((size_t)nb);
  std::vector<uint32_t> ro((size_t)nb), rp((size_t)std::max<int64_t>(d.nref_row_ptr, 1));
  std::vector<uint16_t> ci((size_t)std::max<int64_t>(nz, 1));
  std::vector<V> val((size_t)std::max<int64_t>(nz, 1));
  const int vb = M.dtype == RVVLA_F64 ? 8 : M.dtype == RVVLA_F32 ? 4 : 2;
  std::vector<uint16_t> v16(vb == 2 ? (size_t)std::max<int64_t>(nz, 1) : 0);
  auto get = [&](void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    if (int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "copy back"))
      throw Error(rc, last_error());
  };
  get(bp.data(), d.ref_block_ptr, (size_t)nb * 4);
  get(ro.data(), d.ref_row_offset, (size_t)nb * 4);
  get(rp.data(), d.ref_row_ptr, (size_t)d.nref_row_ptr * 4);
  get(ci.data(), d.ref_col_idx, (size_t)nz * 2);
  get(vb == 2 ? (void*)v16.data() : (void*)val.data(), d.ref_values, (size_t)nz * vb);
  if (int rc = hip_check(hipStreamSynchronize(s), "copy back")) throw Error(rc, last_error());
  if (vb == 2)
    for (int64_t i = 0; i < nz; ++i)
      val[(size_t)i] = (V)(M.dtype == RVVLA_BF16 ? bf16_to_f32(v16[(size_t)i]) : f16_to_f32(v16[(size_t)i]));
  HcsrImage<V> im;
  build_hcsr_image(im, M.m, M.n, d.R, d.C, bp.data(), ro.data(), rp.data(), ci.data(), val.data());
  put_hcsr_image(M, M.dtype, im);
}
template void hcsr_image_from_device<double>(rvvla_mat_s&, hipStream_t);
template void hcsr_image_from_device<float>(rvvla_mat_s&, hipStream_t);

template <class V>
rvvla_mat_t upload_hcsr(int dtype, const spMtxHCSR<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t m = (int64_t)a.m, n = (int64_t)a.n, nz = (int64_t)a.nz;
  const int64_t R = (int64_t)a.R, C = (int64_t)a.C;
  if (R <= 0 || C <= 0 || C > 65536) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: bad R or C");
  if (nz > 2147483647LL || m > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "HCSR: too large");
  const int64_t bm = (m + R - 1) / R, bn = (n + C - 1) / C;
  if ((int64_t)a.b_m != bm || (int64_t)a.b_n != bn)
    throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: b_m/b_n do not match m, n, R, C");
  const size_t nblocks = (size_t)(bm * bn);
  if (a.block_ptr.size() != nblocks + 1 || a.row_offset.size() != nblocks + 1 ||
      a.col_idx.size() != (size_t)nz || a.values.size() != (size_t)nz)
    throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: array sizes do not match the block grid");
  if (a.block_ptr[0] != 0 || a.block_ptr[nblocks] != nz)
    throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: block_ptr must run from 0 to nz");
  for (int64_t br = 0; br < bm; ++br) {
    const int64_t rows = std::min<int64_t>(R, m - br * R);
    for (int64_t bc = 0; bc < bn; ++bc) {
      const size_t b = (size_t)(br * bn + bc);
      const int64_t e0 = a.block_ptr[b], e1 = a.block_ptr[b + 1];
      if (e1 < e0) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: block_ptr not monotone");
      const uint32_t kind = a.row_offset[b] >> 30, off = a.row_offset[b] & 0x3FFFFFFFu;
      const int64_t cols_here = std::min<int64_t>(C, n - bc * C);
      for (int64_t e = e0; e < e1; ++e)
        if (a.col_idx[(size_t)e] >= cols_here)
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: col_idx outside its block");
      if (kind == 0) {
        if (e1 != e0) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: empty-kind block holds entries");
      } else if (kind == 1) {
        if ((size_t)off + (size_t)rows + 1 > a.row_ptr.size())
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: row_ptr too short");
        const uint32_t* rp = a.row_ptr.data() + off;
        if (rp[0] != 0 || (int64_t)rp[rows] != e1 - e0)
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: local row_ptr does not span the block");
        for (int64_t r = 0; r < rows; ++r)
          if (rp[r + 1] < rp[r]) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: local row_ptr");
      } else if (kind == 2) {
        if ((size_t)off + (size_t)(e1 - e0) > a.row_ptr.size())
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: row_ptr too short");
        for (int64_t e = e0; e < e1; ++e)
          if ((int64_t)a.row_ptr[off + (size_t)(e - e0)] >= rows)
            throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: COO row out of block");
      } else {
        throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: unknown block kind");
      }
    }
  }
  HcsrImage<V> im;
  build_hcsr_image(im, m, n, R, C, a.block_ptr.data(), a.row_offset.data(), a.row_ptr.data(),
                   a.col_idx.data(), a.values.data());
  auto M = new_mat(RVVLA_FMT_HCSR, dtype, m, n, nz);
  DevHCSR& d = M->hcsr;
  d.R = R;
  d.C = C;
  d.b_m = bm;
  d.b_n = bn;
  put_hcsr_image(*M, dtype, im);
  return M.release();
}

// Pieces of the slices wider than kSellPiece steps (see DevSELL).
void attach_sell_pieces(rvvla_mat_s& M, const std::vector<int>& cl) {
  std::vector<int32_t> pc_slice, pc_j0, pc_j1, sp_slice, sp_first{0};
  for (size_t i = 0; i < cl.size(); ++i) {
    if (cl[i] <= kSellPiece) continue;
    for (int j = 0; j < cl[i]; j += kSellPiece) {
      pc_slice.push_back((int32_t)i);
      pc_j0.push_back(j);
      pc_j1.push_back(std::min(cl[i], j + kSellPiece));
    }
    sp_slice.push_back((int32_t)i);
    sp_first.push_back((int32_t)pc_slice.size());
  }
  DevSELL& d = M.sell;
  d.npieces = (int64_t)pc_slice.size();
  d.nsplit = (int64_t)sp_slice.size();
  if (d.nsplit == 0) return;
  d.pc_slice = M.put_vec(pc_slice);
  d.pc_j0 = M.put_vec(pc_j0);
  d.pc_j1 = M.put_vec(pc_j1);
  d.sp_slice = M.put_vec(sp_slice);
  d.sp_first = M.put_vec(sp_first);
  d.part = M.put(nullptr, (size_t)d.npieces * 64 * (M.dtype == RVVLA_F64 ? 8 : 4));
}

template <class V>
rvvla_mat_t upload_sell(int dtype, const spMtxSELL_C_Sigma<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t m = (int64_t)a.m, n = (int64_t)a.n, nb = (int64_t)a.cnt_b;
  if (a.C < 1 || a.C > 64) throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "SELL: C must be in [1, 64]");
  if (m > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "SELL: too many rows");
  if ((int64_t)a.vCl.size() != nb || (int64_t)a.vCs.size() != nb || (int64_t)a.vBs.size() != nb ||
      (int64_t)a.vSBs.size() != nb || (int64_t)a.vPerm.size() != m ||
      a.vCol.size() != a.vVal.size())
    throw Error(RVVLA_ERR_INVALID_ARG, "SELL: array sizes do not match");
  const int64_t slots = (int64_t)a.vCol.size();
  int64_t rows_covered = 0;
  for (int64_t i = 0; i < nb; ++i) {
    if (a.vBs[i] < 0 || a.vBs[i] > a.C || a.vCl[i] < 0 || a.vCs[i] < 0 || a.vSBs[i] < 0 ||
        (int64_t)a.vSBs[i] + a.vBs[i] > m ||
        (int64_t)a.vCs[i] + (int64_t)a.vCl[i] * a.vBs[i] > slots)
      throw Error(RVVLA_ERR_INVALID_ARG, "SELL: slice descriptor out of range");
    rows_covered += a.vBs[i];
  }
  if (rows_covered != m) throw Error(RVVLA_ERR_INVALID_ARG, "SELL: slices do not cover m rows");
  for (int64_t s = 0; s < slots; ++s)
    if ((int64_t)a.vCol[(size_t)s] >= n) throw Error(RVVLA_ERR_INVALID_ARG, "SELL: column out of range");
  {
    std::vector<char> seen((size_t)m, 0);
    for (int64_t s = 0; s < m; ++s) {
      const uint32_t r = a.vPerm[(size_t)s];
      if ((int64_t)r >= m || seen[r]) throw Error(RVVLA_ERR_INVALID_ARG, "SELL: vPerm is not a permutation");
      seen[r] = 1;
    }
  }
  auto M = new_mat(RVVLA_FMT_SELL, dtype, m, n, (int64_t)a.nz);
  fill_sell(*M, dtype, a);
  return M.release();
}

// The SELL-C-sigma device image of a checked container into M.sell.
template <class V>
static void fill_sell(rvvla_mat_s& Mr, int dtype, const spMtxSELL_C_Sigma<V>& a,
                      bool keep_perm) {
  const int64_t m = (int64_t)a.m, nb = (int64_t)a.cnt_b;
  rvvla_mat_s* M = &Mr;
  DevSELL& d = M->sell;
  d.nslices = nb;
  d.cl = M->put_vec(a.vCl);
  d.cs = M->put_vec(a.vCs);
  d.bs = M->put_vec(a.vBs);
  d.sbs = M->put_vec(a.vSBs);
  d.col = M->put_vec(a.vCol);
  d.perm_identity = true;
  for (int64_t s = 0; s < m && d.perm_identity; ++s) d.perm_identity = a.vPerm[(size_t)s] == (uint32_t)s;
  // the identity permutation is never read by the kernels; the CSR image
  // does not keep it
  d.perm = d.perm_identity && !keep_perm ? nullptr : M->put_vec(a.vPerm);
  d.val = put_values(*M, dtype, a.vVal.data(), a.vVal.size());
  attach_sell_pieces(*M, a.vCl);
}

static std::vector<int32_t> empty_row_list(int64_t m, const int* rst) {
  std::vector<int32_t> e;
  for (int64_t r = 0; r < m; ++r)
    if (rst[r] == rst[r + 1]) e.push_back((int32_t)r);
  return e;
}

template <class V>
rvvla_mat_t upload_csr5(int dtype, const spMtxCSR5<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  if (a.format() != CSR5_FORMAT_CSR5) throw Error(RVVLA_ERR_UNKNOWN_FORMAT, "CSR5: not converted");
  if (a._omega != 64) throw Error(RVVLA_ERR_UNSUPPORTED_OMEGA, "CSR5: omega must be 64");
  if (a._sigma < 1 || a._sigma > 64 || a._tile_desc_len < 1 || a._tile_desc_len > 3)
    throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "CSR5: sigma must be in [1, 64]");
  if (a._bit_y_offset < 1 || a._bit_seg_offset < 0 || a._bit_y_offset + a._bit_seg_offset > 31 ||
      a._bit_y_offset + a._bit_seg_offset + a._sigma > 32 * a._tile_desc_len)
    throw Error(RVVLA_ERR_INVALID_ARG, "CSR5: descriptor bit fields do not fit the descriptor");
  const int64_t m = a._m, n = a._n, nnz = a._nnz, W = 64LL * a._sigma, ntiles = a._p - 1;
  check_csr(m, n, nnz, a._csr_row_pointer.data(), (const int*)a._csr_column_index.data());
  if ((int64_t)a._csr5_tile_pointer.size() != a._p + 1 ||
      (int64_t)a._csr5_tile_descriptor.size() != std::max<int64_t>(ntiles, 0) * 64 * a._tile_desc_len ||
      ntiles * W > nnz)
    throw Error(RVVLA_ERR_INVALID_ARG, "CSR5: array sizes do not match");
  for (int64_t t = 0; t <= a._p; ++t)
    if ((int64_t)(a._csr5_tile_pointer[(size_t)t] & 0x7FFFFFFFu) > m)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR5: tile_ptr out of range");
  const int* rst = a._csr_row_pointer.data();
  auto M = new_mat(RVVLA_FMT_CSR5, dtype, m, n, nnz);
  DevCSR5& d = M->csr5;
  d.ntiles = std::max<int64_t>(ntiles, 0);
  d.sigma = a._sigma;
  d.dl = a._tile_desc_len;
  d.bit_y = a._bit_y_offset;
  d.bit_seg = a._bit_seg_offset;
  d.tail_base = d.ntiles * W;
  d.tail_start = a._tail_tile_start;
  d.first_tail_row = (a._tail_tile_start < m && rst[a._tail_tile_start] < d.tail_base)
                         ? a._tail_tile_start + 1 : a._tail_tile_start;
  const std::vector<int32_t> empty = empty_row_list(m, rst);
  d.nempty = (int64_t)empty.size();
  d.tile_ptr = M->put_vec(a._csr5_tile_pointer);
  d.desc = M->put_vec(a._csr5_tile_descriptor);
  d.eptr = M->put_vec(a._csr5_tile_empty_offset_pointer);
  d.eoff = M->put_vec(a._csr5_tile_empty_offset);
  d.rst = M->put_vec(a._csr_row_pointer);
  d.col = M->put_vec(a._csr_column_index);
  d.val = put_values(*M, dtype, a._csr_value.data(), a._csr_value.size());
  d.empty_rows = M->put_vec(empty);
  d.np = a._p;
  d.noffsets = (int64_t)a._csr5_tile_empty_offset.size();
  // Where each open segment ends, from the descriptor flags (the same bits
  // the tile kernel reads): bit 0 of kind = a row starts at the tile's first
  // entry, bit 1 = a row starts later in the tile.  A segment left open by
  // tile t runs through the following tiles up to the first one with a row
  // start, including that tile's leading segment unless the start is its
  // first entry; with no such tile it runs into the tail.
  {
    const int f0 = 32 - a._bit_y_offset - a._bit_seg_offset, dl = a._tile_desc_len, sg = a._sigma;
    const uint32_t* desc = a._csr5_tile_descriptor.data();
    std::vector<uint8_t> kind((size_t)d.ntiles);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < d.ntiles; ++t) {
      bool first = false, later = false;
      for (int l = 0; l < 64; ++l) {
        const uint32_t* w = desc + t * 64 * dl + l;
        for (int i = 0; i < sg; ++i) {
          const uint32_t bit = i < f0 ? (w[0] >> (f0 - 1 - i)) & 1u
                                      : (w[64 * (1 + (i - f0) / 32)] >> (31 - (i - f0) % 32)) & 1u;
          if (!bit) continue;
          if (l == 0 && i == 0) first = true;
          else later = true;
        }
      }
      kind[(size_t)t] = (uint8_t)((first ? 1 : 0) | (later ? 2 : 0));
    }
    std::vector<int32_t> chain((size_t)d.ntiles), longs;
    int64_t next = d.ntiles;  // the first tile after t with a row start
    for (int64_t t = d.ntiles - 1; t >= 0; --t) {
      const uint32_t end = next == d.ntiles ? ((uint32_t)d.ntiles | 0x80000000u)
                           : (kind[(size_t)next] & 1) ? (uint32_t)next : (uint32_t)next + 1;
      chain[(size_t)t] = (int32_t)end;
      if (kind[(size_t)t] != 0 &&
          ((end >> 31) || (int64_t)(end & 0x7FFFFFFFu) - t - 1 > kCsr5Serial))
        longs.push_back((int32_t)t);
      if (kind[(size_t)t] != 0) next = t;
    }
    for (int64_t r = d.first_tail_row; r < m; ++r)
      if (rst[r + 1] - rst[r] > kCsr5Serial) longs.push_back((int32_t)(-(r + 1)));
    d.chain = M->put_vec(chain);
    d.nlong = (int64_t)longs.size();
    d.long_fix = M->put_vec(longs);
  }
  const size_t acc = dtype == RVVLA_F64 ? 8 : 4;
  d.scratch = M->put(nullptr, (size_t)d.ntiles * (2 * acc + 4) + 16);
  M->device_bytes -= (int64_t)((size_t)d.ntiles * (2 * acc + 4) + 16);
  return M.release();
}

template <class V>
rvvla_mat_t upload_cvr(int dtype, const spMtxCVR<V>& a, int64_t nnz) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t L = (int64_t)a.SIMD_LEN, m = (int64_t)a.numRows, n = (int64_t)a.numCols;
  const int64_t S = a.Nthreads;
  if (L != 64) throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "CVR: SIMD_LEN must be 64 on gfx950");
  if (S < 1 || (int64_t)a.writeback_pos_record.size() != S * (L + 2) ||
      a.vPack_vec_vals.size() != a.nItems || a.vPack_vec_cols.size() != a.nItems ||
      a.nItems % L != 0 || a.vPack_vec_record.size() % 2 != 0 || a.vPack_vec_record.size() < 2)
    throw Error(RVVLA_ERR_INVALID_ARG, "CVR: array sizes do not match");
  for (uint32_t c : a.vPack_vec_cols)
    if ((int64_t)c >= n) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: column out of range");
  const int64_t npacks = (int64_t)a.nItems / L, nrec = (int64_t)a.vPack_vec_record.size() / 2 - 1;
  std::vector<int64_t> pstart((size_t)S + 1);
  std::vector<int32_t> rstart((size_t)S), rec_rows, tails((size_t)(S * 64), -1);
  std::vector<uint64_t> flush((size_t)npacks, 0);
  std::vector<int32_t> times((size_t)m, 0);  // records of each row
  const int* wb = a.writeback_pos_record.data();
  for (int64_t t = 0; t <= S; ++t) pstart[(size_t)t] = t < S ? wb[t * (L + 2)] : npacks;
  auto rec_range = [&](int64_t t) {
    return std::make_pair((int64_t)wb[t * (L + 2) + 1], t + 1 < S ? (int64_t)wb[(t + 1) * (L + 2) + 1] : nrec);
  };
  for (int64_t t = 0; t < S; ++t) {
    if (pstart[(size_t)t] < 0 || pstart[(size_t)t] > pstart[(size_t)t + 1])
      throw Error(RVVLA_ERR_INVALID_ARG, "CVR: stream pack ranges");
    const auto [r0, r1] = rec_range(t);
    if (r0 < 0 || r0 > r1 || r1 > nrec) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: record ranges");
    int64_t prev = -1;
    for (int64_t q = r0; q < r1; ++q) {
      const int64_t pos = a.vPack_vec_record[(size_t)(2 * q)];
      const int64_t row = a.vPack_vec_record[(size_t)(2 * q + 1)];
      const int64_t k = pos / L, lane = pos % L;
      if (row < 0 || row >= m || k < pstart[(size_t)t] || k > pstart[(size_t)t + 1])
        throw Error(RVVLA_ERR_INVALID_ARG, "CVR: record out of range");
      // records must come in (pack, lane) order inside a stream for the
      // rank lookup
      if (pos <= prev) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: records not in position order");
      prev = pos;
      ++times[(size_t)row];
    }
  }
  // rows recorded more than once: their slots, row by row (each row's in
  // record order); listed for the combine with the rows of more than 64
  // slots first, then the others, each part in row order
  std::vector<int32_t> split_rows, split_ptr, sbase((size_t)m, -1);
  int64_t nslots = 0;
  for (int64_t r = 0; r < m; ++r)
    if (times[(size_t)r] > 1) {
      sbase[(size_t)r] = (int32_t)nslots;
      nslots += times[(size_t)r];
    }
  if (nslots > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "CVR: too many partial sums");
  for (int pass = 0; pass < 2; ++pass)
    for (int64_t r = 0; r < m; ++r)
      if (times[(size_t)r] > 1 && (times[(size_t)r] > 64) == (pass == 0)) {
        split_rows.push_back((int32_t)r);
        split_ptr.push_back(sbase[(size_t)r]);
        split_ptr.push_back(sbase[(size_t)r] + times[(size_t)r]);
      }
  int64_t nsplit_long = 0;
  for (int64_t r = 0; r < m; ++r) nsplit_long += times[(size_t)r] > 64;
  std::vector<int32_t> seen((size_t)m, 0);
  for (int64_t t = 0; t < S; ++t) {
    const auto [r0, r1] = rec_range(t);
    rstart[(size_t)t] = (int32_t)rec_rows.size();
    for (int64_t q = r0; q < r1; ++q) {
      const int64_t pos = a.vPack_vec_record[(size_t)(2 * q)];
      const int64_t row = a.vPack_vec_record[(size_t)(2 * q + 1)];
      const int64_t k = pos / L, lane = pos % L;
      const int32_t target = sbase[(size_t)row] < 0
          ? (int32_t)row : cvr_slot_target(sbase[(size_t)row] + seen[(size_t)row]++);
      if (k == pstart[(size_t)t + 1]) {
        tails[(size_t)(t * 64 + lane)] = target;  // flushed after the stream's last pack
      } else {
        if ((flush[(size_t)k] >> lane) & 1u) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: duplicate record");
        flush[(size_t)k] |= uint64_t(1) << lane;
        rec_rows.push_back(target);
      }
    }
  }
  std::vector<int32_t> empty;
  for (int64_t r = 0; r < m; ++r)
    if (!times[(size_t)r]) empty.push_back((int32_t)r);
  auto M = new_mat(RVVLA_FMT_CVR, dtype, m, n, nnz);
  DevCVR& d = M->cvr;
  d.nstreams = S;
  d.nempty = (int64_t)empty.size();
  d.npacks = npacks;
  d.nrec = (int64_t)rec_rows.size();
  d.pack_start = M->put_vec(pstart);
  d.rec_start = M->put_vec(rstart);
  d.flush = M->put_vec(flush);
  d.rec_rows = M->put_vec(rec_rows);
  d.tail_rows = M->put_vec(tails);
  d.col = M->put_vec(a.vPack_vec_cols);
  d.val = put_values(*M, dtype, a.vPack_vec_vals.data(), a.vPack_vec_vals.size());
  d.empty_rows = M->put_vec(empty);
  d.nsplit = (int64_t)split_rows.size();
  d.nsplit_long = nsplit_long;
  d.nslots = nslots;
  if (d.nsplit > 0) {
    d.split_rows = M->put_vec(split_rows);
    d.split_ptr = M->put_vec(split_ptr);
    d.part = M->put(nullptr, (size_t)d.nslots * (dtype == RVVLA_F64 ? 8 : 4));
  }
  return M.release();
}

template <class V>
rvvla_mat_t upload_vhcc(int dtype, const spMtxVHCC<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const vhcc_params P = a.get_params();
  const int64_t m = (int64_t)a.get_num_rows(), n = (int64_t)a.get_num_cols();
  const int64_t nvec = (int64_t)a.get_num_vectors(), T = (int64_t)a.thr_info().size();
  if (P.vec_size != 64) throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "VHCC: vec_size must be 64 on gfx950");
  const int nm = (int)P.n_masks, mb = 8;
  if (nm != 6 || (int64_t)a.vals_arr().size() != nvec * 64 || a.col_arr().size() != a.vals_arr().size() ||
      T < 1 || a.veceor().empty() || a.scan_mask().size() != (a.veceor().size() - 1) * (size_t)(nm + 1) * mb)
    throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: array sizes do not match");
  for (uint32_t c : a.col_arr())
    if ((int64_t)c >= n) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: column out of range");
  std::vector<uint64_t> eor((size_t)nvec, 0);
  const size_t neor = a.veceor().size() - 1;
  size_t nends = 0;
  for (size_t k = 0; k < neor; ++k) {
    const int64_t v = a.veceor()[k];
    if (v < 0 || v >= nvec || (k > 0 && a.veceor()[k - 1] >= (uint32_t)v))
      throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: row-end vector list");
    uint64_t w = 0;
    std::memcpy(&w, a.scan_mask().data() + (k * (nm + 1) + nm) * mb, 8);
    if (!w) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: empty write mask");
    eor[(size_t)v] = w;
    nends += (size_t)__builtin_popcountll(w);
  }
  if (nends != a.row_arr().size()) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: row_arr length");
  std::vector<int32_t> sv((size_t)T + 1), rb((size_t)T), fwd((size_t)T, -1);
  std::vector<uint8_t> cont((size_t)T), pass((size_t)T, 0);
  int64_t ends_before = 0, vprev = 0;
  for (int64_t t = 0; t < T; ++t) {
    const thr_info_t& I = a.thr_info()[(size_t)t];
    if (I.start_vec != vprev || I.end_vec < I.start_vec || I.end_vec > nvec)
      throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: partitions must tile the vectors in order");
    sv[(size_t)t] = I.start_vec;
    rb[(size_t)t] = (int32_t)ends_before;
    for (int64_t v = I.start_vec; v < I.end_vec; ++v) ends_before += __builtin_popcountll(eor[(size_t)v]);
    vprev = I.end_vec;
    cont[(size_t)t] = I.last_row != I.overflow_row;
    if (cont[(size_t)t] && (I.last_row < 0 || I.last_row >= m))
      throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: last_row out of range");
  }
  if (vprev != nvec) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: partitions do not cover the vectors");
  sv[(size_t)T] = (int32_t)nvec;
  for (uint32_t r : a.row_arr())
    if ((int64_t)r >= m + T) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: row_arr out of range");
  for (int64_t t = 0; t < T; ++t) {
    const bool nxt = t + 1 < T && cont[(size_t)t + 1];
    const int nrow = nxt ? a.thr_info()[(size_t)t + 1].last_row : -1;
    pass[(size_t)t] = cont[(size_t)t] && nxt && a.thr_info()[(size_t)t].last_row == nrow;
    fwd[(size_t)t] = (nxt && !pass[(size_t)t]) ? nrow : -1;
  }
  // rows that never receive a value: empty rows
  std::vector<char> hit((size_t)m, 0);
  for (uint32_t r : a.row_arr())
    if ((int64_t)r < m) hit[r] = 1;
  std::vector<int32_t> empty;
  for (int64_t r = 0; r < m; ++r)
    if (!hit[(size_t)r]) empty.push_back((int32_t)r);
  auto M = new_mat(RVVLA_FMT_VHCC, dtype, m, n, (int64_t)a.get_num_entries());
  DevVHCC& d = M->vhcc;
  d.nparts = T;
  d.nempty = (int64_t)empty.size();
  d.nvec = nvec;
  d.nends = (int64_t)nends;
  d.start_vec = M->put_vec(sv);
  d.rbase = M->put_vec(rb);
  d.eor = M->put_vec(eor);
  d.row_arr = M->put_vec(a.row_arr());
  d.fwd_row = M->put_vec(fwd);
  d.pass = M->put_vec(pass);
  d.col = M->put_vec(a.col_arr());
  d.val = put_values(*M, dtype, a.vals_arr().data(), a.vals_arr().size());
  d.empty_rows = M->put_vec(empty);
  const size_t acc = dtype == RVVLA_F64 ? 8 : 4;
  d.scratch = M->put(nullptr, (size_t)T * 2 * acc + 16);
  M->device_bytes -= (int64_t)((size_t)T * 2 * acc + 16);
  return M.release();
}

// Chunks with more than kLavPiece steps over all segments -> pieces of at
// most kLavPiece steps (within one segment), from the chunk offsets
// (S * (cc + 1), per segment) and the segment bases (in steps).
static void attach_lav_pieces(rvvla_mat_s& M, int64_t S, int64_t cc, const int* coff,
                              const int64_t* sptr) {
  DevLAV& d = M.lav;
  std::vector<int32_t> whole, pc_chunk, sp_chunk, sp_first{0};
  std::vector<int64_t> pc_k0, pc_k1;
  for (int64_t c = 0; c < cc; ++c) {
    int64_t steps = 0;
    for (int64_t g = 0; g < S; ++g)
      steps += coff[(size_t)(g * (cc + 1) + c + 1)] - coff[(size_t)(g * (cc + 1) + c)];
    if (steps <= kLavPiece) {
      whole.push_back((int32_t)c);
      continue;
    }
    for (int64_t g = 0; g < S; ++g) {
      const int64_t k0 = sptr[(size_t)g] + coff[(size_t)(g * (cc + 1) + c)];
      const int64_t k1 = sptr[(size_t)g] + coff[(size_t)(g * (cc + 1) + c + 1)];
      for (int64_t k = k0; k < k1; k += kLavPiece) {
        pc_chunk.push_back((int32_t)c);
        pc_k0.push_back(k);
        pc_k1.push_back(std::min(k1, k + kLavPiece));
      }
    }
    sp_chunk.push_back((int32_t)c);
    sp_first.push_back((int32_t)pc_chunk.size());
  }
  d.nwhole = (int64_t)whole.size();
  if (!sp_chunk.empty()) {
    d.whole = M.put_vec(whole);
    d.npieces = (int64_t)pc_chunk.size();
    d.nsplit = (int64_t)sp_chunk.size();
    d.pc_chunk = M.put_vec(pc_chunk);
    d.pc_k0 = M.put_vec(pc_k0);
    d.pc_k1 = M.put_vec(pc_k1);
    d.sp_chunk = M.put_vec(sp_chunk);
    d.sp_first = M.put_vec(sp_first);
    d.part = M.put(nullptr, (size_t)d.npieces * 64 * (M.dtype == RVVLA_F64 ? 8 : 4));
  }
}

template <class V>
rvvla_mat_t upload_lav(int dtype, const spMtxLAV<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t L = (int64_t)a.SIMD_Lanes, m = a.m, n = a.n, S = (int64_t)a.segment_count;
  const int64_t cc = (int64_t)a.chunk_count;
  if (L != 64) throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "LAV: SIMD_Lanes must be 64 on gfx950");
  const spMtxCRS<V>& sp = a.sparse_part;
  if ((int64_t)sp.m != m || (int64_t)sp.n != n || cc != (m + L - 1) / L ||
      (int64_t)a.chunk_offsets.size() != S * (cc + 1) ||
      (int64_t)a.chunk_offsets_full_mask.size() != S * cc || (int64_t)a.segment_ptr.size() < S ||
      (int64_t)a.out_order.size() != S * cc * L || a.val.size() != a.col_id.size() ||
      a.val.size() != a.mask.size() * (size_t)L)
    throw Error(RVVLA_ERR_INVALID_ARG, "LAV: array sizes do not match");
  check_csr(m, n, (int64_t)sp.nz, sp.Rst, sp.Col);
  const int64_t total = (int64_t)a.mask.size();
  for (uint32_t c : a.col_id)
    if ((int64_t)c >= n) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: column out of range");
  if (S > 0) {  // no dense part: out_order is empty and never read
    std::vector<char> seen((size_t)m, 0);
    for (int64_t i = 0; i < m; ++i) {
      const uint32_t r = a.out_order[(size_t)i];
      if ((int64_t)r >= m || seen[r]) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: out_order is not a row permutation");
      seen[r] = 1;
    }
    for (int64_t i = m; i < cc * L; ++i)
      if ((int64_t)a.out_order[(size_t)i] >= m) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: out_order out of range");
    for (int64_t g = 1; g < S; ++g)
      if (!std::equal(a.out_order.begin(), a.out_order.begin() + cc * L,
                      a.out_order.begin() + g * cc * L))
        throw Error(RVVLA_ERR_UNSUPPORTED_SPMV,
                    "LAV: the gfx950 kernel needs the same row order in every segment");
  }
  std::vector<int64_t> sptr((size_t)S);
  std::vector<V> vals(a.val);
  // lanes of each step: 64 up to the chunk's full-mask offset, the step's
  // mask after it (the kernels drop the products of the other lanes, so a
  // padding slot never multiplies x, not even an Inf)
  std::vector<uint16_t> lanes((size_t)total, (uint16_t)L);
  for (int64_t g = 0; g < S; ++g) {
    sptr[(size_t)g] = (int64_t)a.segment_ptr[(size_t)g];
    for (int64_t c = 0; c < cc; ++c) {
      const int64_t k0 = a.chunk_offsets[(size_t)(g * (cc + 1) + c)];
      const int64_t k1 = a.chunk_offsets[(size_t)(g * (cc + 1) + c + 1)];
      const int64_t kf = a.chunk_offsets_full_mask[(size_t)(g * cc + c)];
      if (k0 < 0 || k1 < k0 || kf < k0 || kf > k1 || sptr[(size_t)g] + k1 > total)
        throw Error(RVVLA_ERR_INVALID_ARG, "LAV: chunk offsets out of range");
      for (int64_t k = kf; k < k1; ++k) {
        const int64_t step = sptr[(size_t)g] + k;
        if (a.mask[(size_t)step] > L) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: mask out of range");
        lanes[(size_t)step] = (uint16_t)a.mask[(size_t)step];
        for (int64_t i = a.mask[(size_t)step]; i < L; ++i) vals[(size_t)(step * L + i)] = V(0);
      }
    }
  }
  auto M = new_mat(RVVLA_FMT_LAV, dtype, m, n, a.nz);
  const std::vector<int32_t> blocks = csr_row_blocks(sp.Rst, 0, m);
  M->csr.rst = (const int32_t*)M->put(sp.Rst, (sp.m + 1) * sizeof(int));
  M->csr.col = (const int32_t*)M->put(sp.Col, sp.nz * sizeof(int));
  M->csr.val = put_values(*M, dtype, sp.Val, sp.nz);
  M->csr.blocks = M->put_vec(blocks);
  M->csr.nblocks = (int64_t)blocks.size() / 2;
  DevLAV& d = M->lav;
  d.nseg = S;
  d.nchunks = cc;
  {
    int64_t longest = 0;
    for (int64_t r = 0; r < m; ++r) longest = std::max<int64_t>(longest, sp.Rst[r + 1] - sp.Rst[r]);
    d.fused = S > 0 && longest <= kLavFuseRow;
  }
  d.coff = M->put_vec(a.chunk_offsets);
  d.sptr = M->put_vec(sptr);
  d.out_order = (const uint32_t*)M->put(S > 0 ? a.out_order.data() : nullptr,
                                        S > 0 ? (size_t)(cc * L) * 4 : 0);
  d.col = M->put_vec(a.col_id);
  d.val = put_values(*M, dtype, vals.data(), vals.size());
  d.mask = M->put_vec(lanes);
  d.cfull = M->put_vec(a.chunk_offsets_full_mask);
  attach_lav_pieces(*M, S, cc, a.chunk_offsets.data(), sptr.data());
  return M.release();
}

template <class V>
rvvla_mat_t upload_vnec(int dtype, const spMtxVNEC<V>& a, const spMtxCRS<V>& src) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t m = (int64_t)src.m, n = (int64_t)src.n, nnz = (int64_t)src.nz;
  const int64_t P = a.NUM_THREADS_VNEC;
  check_csr(m, n, nnz, src.Rst, src.Col);
  if (P < 1 || (int64_t)a.thread_coord_start.size() != P || (int64_t)a.thread_coord_end.size() != P ||
      (int64_t)a.nec_ptr.size() != P + 1 || (int64_t)a.ecr_indices.size() != nnz ||
      (int64_t)a.use_x_indices.size() != a.nec_ptr[(size_t)P] || a.m != m || a.nnz != nnz)
    throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: array sizes do not match");
  const int64_t items = vnec_items(dtype == RVVLA_F64 ? 8 : 4);  // LDS staging per partition
  std::vector<int32_t> cx((size_t)P + 1), cy((size_t)P + 1);
  for (int64_t t = 0; t < P; ++t) {
    cx[(size_t)t] = a.thread_coord_start[(size_t)t].x;
    cy[(size_t)t] = a.thread_coord_start[(size_t)t].y;
    if (t > 0 && (a.thread_coord_start[(size_t)t].x != a.thread_coord_end[(size_t)t - 1].x ||
                  a.thread_coord_start[(size_t)t].y != a.thread_coord_end[(size_t)t - 1].y))
      throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: partitions must be contiguous");
  }
  cx[(size_t)P] = a.thread_coord_end[(size_t)P - 1].x;
  cy[(size_t)P] = a.thread_coord_end[(size_t)P - 1].y;
  if (cx[0] != 0 || cy[0] != 0 || cx[(size_t)P] != m || cy[(size_t)P] != nnz)
    throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: partitions must cover rows and entries");
  std::vector<uint16_t> ecr((size_t)nnz);
  for (int64_t t = 0; t < P; ++t) {
    const int64_t x0 = cx[(size_t)t], x1 = cx[(size_t)t + 1], y0 = cy[(size_t)t], y1 = cy[(size_t)t + 1];
    const int64_t nec = a.nec_ptr[(size_t)t + 1] - a.nec_ptr[(size_t)t];
    if (x1 < x0 || y1 < y0 || y1 - y0 > items || nec < 0 || nec > items ||
        x1 - x0 + 1 > (int64_t)1 << 30)
      throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "VNEC: a partition exceeds the LDS staging size");
    // the merge path: rows before x0 end at or before y0, row x0 has not ended
    if ((x0 > 0 && src.Rst[x0] > y0) || (x0 < m && src.Rst[x0 + 1] < y0 && x1 > x0))
      throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: coordinates are not on the merge path");
    for (int64_t e = y0; e < y1; ++e) {
      const int k = a.ecr_indices[(size_t)e];
      if (k < 0 || k >= nec || a.use_x_indices[(size_t)(a.nec_ptr[(size_t)t] + k)] != src.Col[e])
        throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: compacted column does not match");
      ecr[(size_t)e] = (uint16_t)k;
    }
  }
  for (int c : a.use_x_indices)
    if (c < 0 || c >= n) throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: column out of range");
  auto M = new_mat(RVVLA_FMT_VNEC, dtype, m, n, nnz);
  DevVNEC& d = M->vnec;
  d.nparts = P;
  d.cx = M->put_vec(cx);
  d.cy = M->put_vec(cy);
  d.nec_ptr = M->put_vec(a.nec_ptr);
  d.use_x = M->put_vec(a.use_x_indices);
  d.nlist = (int64_t)a.use_x_indices.size();
  d.ecr = M->put_vec(ecr);
  d.rst = (const int32_t*)M->put(src.Rst, (size_t)(m + 1) * 4);
  d.val = put_values(*M, dtype, src.Val, (size_t)nnz);
  const size_t acc = dtype == RVVLA_F64 ? 8 : 4;
  d.scratch = M->put(nullptr, (size_t)P * 2 * acc + 16);
  M->device_bytes -= (int64_t)((size_t)P * 2 * acc + 16);
  return M.release();
}
template rvvla_mat_t upload_vnec<double>(int, const spMtxVNEC<double>&, const spMtxCRS<double>&);
template rvvla_mat_t upload_vnec<float>(int, const spMtxVNEC<float>&, const spMtxCRS<float>&);

#define RVVLA_INST_UPLOAD(F, T) template rvvla_mat_t F<double>(int, const T<double>&); \
  template rvvla_mat_t F<float>(int, const T<float>&);
template rvvla_mat_t upload_csr<double>(int, const spMtxCRS<double>&, int, int, int, int);
template rvvla_mat_t upload_csr<float>(int, const spMtxCRS<float>&, int, int, int, int);
RVVLA_INST_UPLOAD(upload_hcsr, spMtxHCSR)
RVVLA_INST_UPLOAD(upload_sell, spMtxSELL_C_Sigma)
RVVLA_INST_UPLOAD(upload_csr5, spMtxCSR5)
RVVLA_INST_UPLOAD(upload_vhcc, spMtxVHCC)
RVVLA_INST_UPLOAD(upload_lav, spMtxLAV)
template rvvla_mat_t upload_cvr<double>(int, const spMtxCVR<double>&, int64_t);
template rvvla_mat_t upload_cvr<float>(int, const spMtxCVR<float>&, int64_t);

}  // namespace rvvla

// ---------------------------------------------------------------- layouts
struct ArrayView {
  const void* data;
  int64_t count;
  int elem_bytes;
};

struct rvvla_layout_s {
  int format = -1;
  int dtype = -1;
  std::shared_ptr<void> crs;  // the source matrix (kept for formats that reference it)
  std::shared_ptr<void> obj;  // the format container
  std::map<std::string, ArrayView> arrays;
  std::map<std::string, double> scalars;
  std::function<rvvla_mat_t()> upload;
};

namespace {

template <class T>
void reg(rvvla_layout_s& L, const char* name, const std::vector<T>& v) {
  L.arrays[name] = ArrayView{v.data(), (int64_t)v.size(), (int)sizeof(T)};
}

template <class V>
void convert_into(rvvla_layout_s& L, int format, int dtype, std::shared_ptr<spMtxCRS<V>> crs,
                  const convertParams& P) {
  using namespace rvvla;
  L.crs = crs;
  switch (format) {
    case RVVLA_FMT_CSR: {
      L.obj = crs;
      const spMtxCRS<V>* a = crs.get();
      L.arrays["Rst"] = ArrayView{a->Rst, (int64_t)a->m + 1, 4};
      L.arrays["Col"] = ArrayView{a->Col, (int64_t)a->nz, 4};
      L.arrays["Val"] = ArrayView{a->Val, (int64_t)a->nz, (int)sizeof(V)};
      L.scalars = {{"m", (double)a->m}, {"n", (double)a->n}, {"nz", (double)a->nz}};
      const int xs = param_int(P, "xslices", -1);
      const int xl = param_int(P, "xlds", -1);
      const int sl = param_int(P, "sell", -1);
      const int xc = param_int(P, "xlcomb", 0);
      L.upload = [a, dtype, xs, xl, sl, xc]() { return upload_csr<V>(dtype, *a, xs, xl, sl, xc); };
      break;
    }
    case RVVLA_FMT_HCSR: {
      // reference defaults: R = C = 2000 (FP64), 4000 (FP32), p = 2
      // (reference src/scalar/crs_to_hcsr.cpp:141-143, :158-160)
      const int dflt = std::is_same<V, double>::value ? 2000 : 4000;
      auto h = std::make_shared<spMtxHCSR<V>>();
      build_hcsr<V>(*h, *crs, (size_t)param_int(P, "R", dflt), (size_t)param_int(P, "C", dflt),
                    param_double(P, "p", 2.0));
      L.obj = h;
      reg(L, "block_ptr", h->block_ptr);
      reg(L, "row_offset", h->row_offset);
      reg(L, "row_ptr", h->row_ptr);
      reg(L, "col_idx", h->col_idx);
      reg(L, "values", h->values);
      L.scalars = {{"R", (double)h->R}, {"C", (double)h->C}, {"m", (double)h->m},
                   {"n", (double)h->n}, {"nz", (double)h->nz}, {"b_m", (double)h->b_m},
                   {"b_n", (double)h->b_n}};
      auto* hp = h.get();
      L.upload = [hp, dtype]() { return upload_hcsr<V>(dtype, *hp); };
      break;
    }
    case RVVLA_FMT_SELL: {
      auto s = std::make_shared<spMtxSELL_C_Sigma<V>>();
      build_sell<V>(*s, *crs, param_int(P, "C", 64),
                    param_int(P, "sigma", default_sell_sigma((int64_t)crs->m, (int64_t)crs->nz, crs->Rst)));
      L.obj = s;
      reg(L, "vCl", s->vCl);
      reg(L, "vCs", s->vCs);
      reg(L, "vBs", s->vBs);
      reg(L, "vSBs", s->vSBs);
      reg(L, "vCol", s->vCol);
      reg(L, "vPerm", s->vPerm);
      reg(L, "vVal", s->vVal);
      L.scalars = {{"m", (double)s->m}, {"n", (double)s->n}, {"nz", (double)s->nz},
                   {"cnt_b", (double)s->cnt_b}, {"C", (double)s->C}, {"Sigma", (double)s->Sigma}};
      auto* sp = s.get();
      L.upload = [sp, dtype]() { return upload_sell<V>(dtype, *sp); };
      break;
    }
    case RVVLA_FMT_CSR5: {
      auto c = std::make_shared<spMtxCSR5<V>>();
      build_csr5<V>(*c, *crs, param_int(P, "omega", 64), param_int(P, "sigma", 0));
      L.obj = c;
      reg(L, "_csr_row_pointer", c->_csr_row_pointer);
      reg(L, "_csr_column_index", c->_csr_column_index);
      reg(L, "_csr_value", c->_csr_value);
      reg(L, "_csr5_tile_pointer", c->_csr5_tile_pointer);
      reg(L, "_csr5_tile_descriptor", c->_csr5_tile_descriptor);
      reg(L, "_csr5_tile_empty_offset_pointer", c->_csr5_tile_empty_offset_pointer);
      reg(L, "_csr5_tile_empty_offset", c->_csr5_tile_empty_offset);
      L.scalars = {{"_m", (double)c->_m}, {"_n", (double)c->_n}, {"_nnz", (double)c->_nnz},
                   {"_p", (double)c->_p}, {"_omega", (double)c->_omega},
                   {"_sigma", (double)c->_sigma}, {"_bit_y_offset", (double)c->_bit_y_offset},
                   {"_bit_seg_offset", (double)c->_bit_seg_offset},
                   {"_tile_desc_len", (double)c->_tile_desc_len},
                   {"_tail_tile_start", (double)c->_tail_tile_start},
                   {"_num_offsets", (double)c->_num_offsets}};
      auto* cp = c.get();
      L.upload = [cp, dtype]() { return upload_csr5<V>(dtype, *cp); };
      break;
    }
    case RVVLA_FMT_CVR: {
      auto c = std::make_shared<spMtxCVR<V>>();
      build_cvr<V>(*c, *crs, param_int(P, "SIMD_LEN", 64),
                   param_int(P, "streams", default_parts((int64_t)crs->m, (int64_t)crs->nz, kCvrStreamEntries)));
      L.obj = c;
      reg(L, "vPack_vec_vals", c->vPack_vec_vals);
      reg(L, "vPack_vec_cols", c->vPack_vec_cols);
      reg(L, "vPack_vec_record", c->vPack_vec_record);
      reg(L, "writeback_pos_record", c->writeback_pos_record);
      L.scalars = {{"nItems", (double)c->nItems}, {"numRows", (double)c->numRows},
                   {"numCols", (double)c->numCols}, {"SIMD_LEN", (double)c->SIMD_LEN},
                   {"Nthreads", (double)c->Nthreads}, {"Nthreads_avail", (double)c->Nthreads_avail}};
      auto* cp = c.get();
      const int64_t nz = (int64_t)crs->nz;
      L.upload = [cp, dtype, nz]() { return upload_cvr<V>(dtype, *cp, nz); };
      break;
    }
    case RVVLA_FMT_VHCC: {
      auto c = std::make_shared<spMtxVHCC<V>>();
      build_vhcc<V>(*c, *crs, param_int(P, "vec_size", 64),
                    param_int(P, "threads", default_parts((int64_t)crs->m, (int64_t)crs->nz, kVhccPartEntries)));
      L.obj = c;
      reg(L, "_veceor_ptr", c->veceor());
      reg(L, "_scan_mask", c->scan_mask());
      reg(L, "_row_arr", c->row_arr());
      reg(L, "_col_arr", c->col_arr());
      reg(L, "_vals_arr", c->vals_arr());
      {
        const auto& ti = c->thr_info();
        L.arrays["_thr_info"] = ArrayView{ti.data(), (int64_t)(ti.size() * 10), 4};
      }
      L.scalars = {{"num_rows", (double)c->get_num_rows()}, {"num_cols", (double)c->get_num_cols()},
                   {"num_entries", (double)c->get_num_entries()},
                   {"num_vectors", (double)c->get_num_vectors()},
                   {"pad_rows", (double)c->get_pad_rows()},
                   {"num_threads", (double)c->get_params().num_threads},
                   {"vec_size", (double)c->get_params().vec_size},
                   {"n_masks", (double)c->get_n_masks()}};
      auto* cp = c.get();
      L.upload = [cp, dtype]() { return upload_vhcc<V>(dtype, *cp); };
      break;
    }
    case RVVLA_FMT_LAV: {
      auto c = std::make_shared<spMtxLAV<V>>();
      build_lav<V>(*c, *crs, param_int(P, "SIMD_Lanes", 64), param_double(P, "T", 0.2),
                   param_int(P, "segment_len", 1 << 20));
      L.obj = c;
      reg(L, "chunk_offsets", c->chunk_offsets);
      reg(L, "chunk_offsets_full_mask", c->chunk_offsets_full_mask);
      reg(L, "segment_ptr", c->segment_ptr);
      reg(L, "out_order", c->out_order);
      reg(L, "mask", c->mask);
      reg(L, "val", c->val);
      reg(L, "col_id", c->col_id);
      const spMtxCRS<V>& sp = c->sparse_part;
      L.arrays["sparse_part.Rst"] = ArrayView{sp.Rst, (int64_t)sp.m + 1, 4};
      L.arrays["sparse_part.Col"] = ArrayView{sp.Col, (int64_t)sp.nz, 4};
      L.arrays["sparse_part.Val"] = ArrayView{sp.Val, (int64_t)sp.nz, (int)sizeof(V)};
      L.scalars = {{"SIMD_Lanes", (double)c->SIMD_Lanes}, {"m", (double)c->m}, {"n", (double)c->n},
                   {"nz", (double)c->nz}, {"segment_count", (double)c->segment_count},
                   {"chunk_count", (double)c->chunk_count}, {"T", c->T},
                   {"segment_len", (double)c->segment_len}, {"sparse_nz", (double)sp.nz}};
      auto* cp = c.get();
      L.upload = [cp, dtype]() { return upload_lav<V>(dtype, *cp); };
      break;
    }
    case RVVLA_FMT_VNEC: {
      auto c = std::make_shared<spMtxVNEC<V>>();
      const int64_t items = (int64_t)crs->m + (int64_t)crs->nz;
      const int64_t cap = vnec_items((int)sizeof(V));
      const int dflt = (int)std::max<int64_t>(1, (items + cap - 1) / cap);
      build_vnec<V>(*c, *crs, param_int(P, "threads", dflt));
      L.obj = c;
      reg(L, "ecr_indices", c->ecr_indices);
      reg(L, "NEC_NUM", c->NEC_NUM);
      reg(L, "nec_ptr", c->nec_ptr);
      reg(L, "use_x_indices", c->use_x_indices);
      L.arrays["thread_coord_start"] =
          ArrayView{c->thread_coord_start.data(), (int64_t)c->thread_coord_start.size() * 2, 4};
      L.arrays["thread_coord_end"] =
          ArrayView{c->thread_coord_end.data(), (int64_t)c->thread_coord_end.size() * 2, 4};
      L.scalars = {{"NUM_THREADS_VNEC", (double)c->NUM_THREADS_VNEC}, {"m", (double)c->m},
                   {"n", (double)c->n}, {"nnz", (double)c->nnz}, {"IRD_mat", (double)c->IRD_mat},
                   {"vnec_t", (double)(int)c->vnec_t}};
      auto* cp = c.get();
      auto* src = crs.get();
      L.upload = [cp, src, dtype]() { return upload_vnec<V>(dtype, *cp, *src); };
      break;
    }
    default:
      throw Error(RVVLA_ERR_UNKNOWN_FORMAT, "format not available");
  }
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    rvvla::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    rvvla::set_last_error("host allocation failed");
    return RVVLA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    rvvla::set_last_error(e.what());
    return RVVLA_ERR_LOGIC;
  }
}

}  // namespace

// ---------------------------------------------------------------- extern "C"
extern "C" {

const char* rvvla_strerror(int code) {
  static thread_local std::string msg;
  const char* base = "unknown error";
  switch (code) {
    case RVVLA_OK: base = "success"; break;
    case RVVLA_ERR_UNKNOWN_FORMAT: base = "unknown format"; break;
    case RVVLA_ERR_UNSUPPORTED_OMEGA: base = "unsupported omega"; break;
    case RVVLA_ERR_CONVERSION: base = "conversion failed"; break;
    case RVVLA_ERR_UNSUPPORTED_SPMV: base = "unsupported spmv"; break;
    case RVVLA_ERR_VALUE_TYPE: base = "unsupported value type"; break;
    case RVVLA_ERR_INVALID_ARG: base = "invalid argument"; break;
    case RVVLA_ERR_LENGTH: base = "length error"; break;
    case RVVLA_ERR_LOGIC: base = "logic error"; break;
    case RVVLA_ERR_HIP: base = "HIP runtime error"; break;
    case RVVLA_ERR_NO_DEVICE: base = "no HIP device"; break;
    case RVVLA_ERR_OUT_OF_MEMORY: base = "out of memory"; break;
    case RVVLA_ERR_NOT_FOUND: base = "not found"; break;
  }
  msg = base;
  if (code != RVVLA_OK && !rvvla::last_error().empty()) msg += ": " + rvvla::last_error();
  return msg.c_str();
}

int rvvla_version(void) { return 100; }

const char* rvvla_format_name(int f) {
  static const char* names[] = {"csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr"};
  return (f >= 0 && f < RVVLA_NUM_FORMATS) ? names[f] : nullptr;
}

int rvvla_convert(int format, int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                  const int32_t* col, const void* val, const char* params, rvvla_layout_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (format < 0 || format >= RVVLA_NUM_FORMATS) throw Error(RVVLA_ERR_UNKNOWN_FORMAT, "format");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "dtype");
    rvvla::check_csr(m, n, nnz, rst, col);
    if (nnz > 0 && !val) throw Error(RVVLA_ERR_INVALID_ARG, "null values");
    const convertParams P = rvvla::parse_params(params);
    auto L = std::make_unique<rvvla_layout_s>();
    L->format = format;
    L->dtype = dtype;
    if (dtype == RVVLA_F64) {
      auto a = std::make_shared<spMtxCRS<double>>((size_t)m, (size_t)n, (size_t)nnz, rst, col,
                                                  (const double*)val);
      convert_into<double>(*L, format, dtype, a, P);
    } else {
      auto a = std::make_shared<spMtxCRS<float>>((size_t)m, (size_t)n, (size_t)nnz, rst, col,
                                                 (const float*)val);
      convert_into<float>(*L, format, dtype, a, P);
    }
    *out = L.release();
    return RVVLA_OK;
  });
}

int rvvla_layout_array(rvvla_layout_t L, const char* field, const void** data, int64_t* count,
                       int* elem_bytes) {
  if (!L || !field || !data || !count || !elem_bytes) return RVVLA_ERR_INVALID_ARG;
  auto it = L->arrays.find(field);
  if (it == L->arrays.end()) {
    rvvla::set_last_error(std::string("no array '") + field + "'");
    return RVVLA_ERR_NOT_FOUND;
  }
  *data = it->second.data;
  *count = it->second.count;
  *elem_bytes = it->second.elem_bytes;
  return RVVLA_OK;
}

int rvvla_layout_scalar(rvvla_layout_t L, const char* field, double* value) {
  if (!L || !field || !value) return RVVLA_ERR_INVALID_ARG;
  auto it = L->scalars.find(field);
  if (it == L->scalars.end()) {
    rvvla::set_last_error(std::string("no scalar '") + field + "'");
    return RVVLA_ERR_NOT_FOUND;
  }
  *value = it->second;
  return RVVLA_OK;
}

int rvvla_layout_free(rvvla_layout_t L) {
  delete L;
  return RVVLA_OK;
}

int rvvla_upload(rvvla_layout_t L, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!L || !out) throw Error(RVVLA_ERR_INVALID_ARG, "null argument");
    *out = L->upload();
    if (L->format != RVVLA_FMT_CSR && *out && L->crs) {
      if (L->dtype == RVVLA_F64)
        rvvla::attach_transpose_source<double>(**out, std::static_pointer_cast<const spMtxCRS<double>>(L->crs));
      else
        rvvla::attach_transpose_source<float>(**out, std::static_pointer_cast<const spMtxCRS<float>>(L->crs));
    }
    return RVVLA_OK;
  });
}

int rvvla_csr_upload(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* Rst,
                     const int32_t* Col, const void* Val, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    rvvla::check_csr(m, n, nnz, Rst, Col);
    if (dtype == RVVLA_F64) {
      spMtxCRS<double> a((size_t)m, (size_t)n, (size_t)nnz, Rst, Col, (const double*)Val);
      *out = rvvla::upload_csr<double>(dtype, a);
    } else {
      spMtxCRS<float> a((size_t)m, (size_t)n, (size_t)nnz, Rst, Col, (const float*)Val);
      *out = rvvla::upload_csr<float>(dtype, a);
    }
    return RVVLA_OK;
  });
}

int rvvla_hcsr_upload(int dtype, int64_t m, int64_t n, int64_t nz, int64_t R, int64_t C,
                      int64_t b_m, int64_t b_n, const int32_t* block_ptr,
                      const uint32_t* row_offset, int64_t row_ptr_len, const uint32_t* row_ptr,
                      const uint16_t* col_idx, const void* values, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out || !block_ptr || !row_offset || (row_ptr_len > 0 && !row_ptr) ||
        (nz > 0 && (!col_idx || !values)) || R <= 0 || C <= 0 || b_m < 0 || b_n < 0)
      throw Error(RVVLA_ERR_INVALID_ARG, "HCSR upload: null or negative argument");
    *out = nullptr;
    const size_t nb = (size_t)(b_m * b_n);
    auto fill = [&](auto& h) {
      using V = typename std::decay_t<decltype(h.values)>::value_type;
      h.m = (size_t)m; h.n = (size_t)n; h.nz = (size_t)nz; h.R = (size_t)R; h.C = (size_t)C;
      h.b_m = (size_t)b_m; h.b_n = (size_t)b_n;
      h.block_ptr.assign(block_ptr, block_ptr + nb + 1);
      h.row_offset.assign(row_offset, row_offset + nb + 1);
      h.row_ptr.assign(row_ptr, row_ptr + row_ptr_len);
      h.col_idx.assign(col_idx, col_idx + nz);
      h.values.assign((const V*)values, (const V*)values + nz);
    };
    if (dtype == RVVLA_F64) {
      spMtxHCSR<double> h;
      fill(h);
      *out = rvvla::upload_hcsr<double>(dtype, h);
    } else {
      spMtxHCSR<float> h;
      fill(h);
      *out = rvvla::upload_hcsr<float>(dtype, h);
    }
    return RVVLA_OK;
  });
}

int rvvla_sell_upload(int dtype, int64_t m, int64_t n, int64_t nz, int64_t cnt_b, int C, int Sigma,
                      const int32_t* vCl, const int32_t* vCs, const int32_t* vBs,
                      const int32_t* vSBs, int64_t slots, const uint32_t* vCol,
                      const uint32_t* vPerm, const void* vVal, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out || cnt_b < 0 || slots < 0 || m < 0 || (cnt_b > 0 && (!vCl || !vCs || !vBs || !vSBs)) ||
        (slots > 0 && (!vCol || !vVal)) || (m > 0 && !vPerm))
      throw Error(RVVLA_ERR_INVALID_ARG, "SELL upload: null or negative argument");
    *out = nullptr;
    auto fill = [&](auto& s) {
      using V = typename std::decay_t<decltype(s.vVal)>::value_type;
      s.m = (size_t)m; s.n = (size_t)n; s.nz = (size_t)nz; s.cnt_b = (size_t)cnt_b;
      s.C = C; s.Sigma = Sigma;
      s.vCl.assign(vCl, vCl + cnt_b); s.vCs.assign(vCs, vCs + cnt_b);
      s.vBs.assign(vBs, vBs + cnt_b); s.vSBs.assign(vSBs, vSBs + cnt_b);
      s.vCol.assign(vCol, vCol + slots); s.vPerm.assign(vPerm, vPerm + m);
      s.vVal.assign((const V*)vVal, (const V*)vVal + slots);
    };
    if (dtype == RVVLA_F64) {
      spMtxSELL_C_Sigma<double> s;
      fill(s);
      *out = rvvla::upload_sell<double>(dtype, s);
    } else {
      spMtxSELL_C_Sigma<float> s;
      fill(s);
      *out = rvvla::upload_sell<float>(dtype, s);
    }
    return RVVLA_OK;
  });
}

int rvvla_sell_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                           const int32_t* d_col, const void* d_val, int sigma, void* stream,
                           rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_SELL, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
                       ? rvvla::build_sell_device<double>(*M, m, nnz, d_rst, d_col, (const double*)d_val, sigma, s)
                       : rvvla::build_sell_device<float>(*M, m, nnz, d_rst, d_col, (const float*)d_val, sigma, s);
    if (rc) throw Error(rc, rvvla::last_error());
    if (M->sell.nslices > 0) {  // pieces of the wide slices, from the slice widths
      std::vector<int> cl((size_t)M->sell.nslices);
      const int rc2 = rvvla::hip_check(
          hipMemcpy(cl.data(), M->sell.cl, cl.size() * 4, hipMemcpyDeviceToHost), "read widths");
      if (rc2) throw Error(rc2, rvvla::last_error());
      rvvla::attach_sell_pieces(*M, cl);
    }
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_lav_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                          const int32_t* d_col, const void* d_val, double T, int64_t segment_len,
                          void* stream, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (T < 0.0) T = 0.2;  // the convert default
    if (segment_len <= 0) segment_len = 1 << 20;
    if (!(T <= 1.0)) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: T in [0, 1]");
    if (segment_len > 2147483647LL) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: segment_len");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_LAV, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    rvvla::LavDeviceHost h;
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_lav_device<double>(*M, m, n, nnz, d_rst, d_col, (const double*)d_val, T, segment_len, h, s)
        : rvvla::build_lav_device<float>(*M, m, n, nnz, d_rst, d_col, (const float*)d_val, T, segment_len, h, s);
    if (rc) throw Error(rc, rvvla::last_error());
    M->free_temps();
    // sparse_part: CSR image (row blocks from its row pointers, on the host)
    std::vector<int32_t> srst((size_t)m + 1);
    int rc2 = rvvla::hip_check(hipMemcpy(srst.data(), h.srst, srst.size() * 4, hipMemcpyDeviceToHost), "read");
    if (rc2) throw Error(rc2, rvvla::last_error());
    const std::vector<int32_t> blocks = rvvla::csr_row_blocks(srst.data(), 0, m);
    M->csr.rst = h.srst;
    M->csr.col = h.scol;
    M->csr.val = h.sval;
    M->csr.blocks = M->put_vec(blocks);
    M->csr.nblocks = (int64_t)blocks.size() / 2;
    rvvla::DevLAV& d = M->lav;
    d.nseg = h.S;
    d.nchunks = h.S > 0 ? h.cc : 0;
    int64_t longest = 0;
    for (int64_t r = 0; r < m; ++r) longest = std::max<int64_t>(longest, srst[(size_t)r + 1] - srst[(size_t)r]);
    d.fused = h.S > 0 && longest <= rvvla::kLavFuseRow;
    d.device_built = true;
    if (h.S == 0) {  // segment_ptr of an empty dense part: one 0
      const int64_t zero = 0;
      d.sptr = (const int64_t*)M->put(&zero, 8);
    } else {
      d.coff = h.coff;
      d.sptr = h.sptr;
      d.out_order = h.out_order;
      d.col = h.col_id;
      d.val = h.dval;
      d.ref_cfull = h.cfull;
      d.ref_mask = h.mask;
      d.mask = h.mask;  // 64 for the full steps, the lane count for the others
      d.cfull = h.cfull;
      d.total_steps = h.total;
      std::vector<int32_t> coff((size_t)(h.S * (h.cc + 1)));
      std::vector<int64_t> sptr((size_t)h.S + 1);
      rc2 = rvvla::hip_check(hipMemcpy(coff.data(), h.coff, coff.size() * 4, hipMemcpyDeviceToHost), "read");
      if (!rc2) rc2 = rvvla::hip_check(hipMemcpy(sptr.data(), h.sptr, sptr.size() * 8, hipMemcpyDeviceToHost), "read");
      if (rc2) throw Error(rc2, rvvla::last_error());
      rvvla::attach_lav_pieces(*M, h.S, h.cc, coff.data(), sptr.data());
    }
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_vnec_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                           const int32_t* d_col, const void* d_val, int64_t nparts, void* stream,
                           rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (nparts <= 0) {  // the default: as many rows + entries per partition as one workgroup stages
      const int64_t cap = rvvla::vnec_items(dtype == RVVLA_F64 ? 8 : 4);
      nparts = std::max<int64_t>(1, (m + nnz + cap - 1) / cap);
    }
    if (nparts > 2147483647LL) throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: too many partitions");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_VNEC, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_vnec_device<double>(*M, m, nnz, d_rst, d_col, (const double*)d_val, nparts, s)
        : rvvla::build_vnec_device<float>(*M, m, nnz, d_rst, d_col, (const float*)d_val, nparts, s);
    if (rc) throw Error(rc, rvvla::last_error());
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_cvr_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                          const int32_t* d_col, const void* d_val, int64_t nstreams, void* stream,
                          rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (nstreams <= 0) nstreams = rvvla::default_parts(m, nnz, rvvla::kCvrStreamEntries);  // the convert default
    if (nstreams > 2147483647LL) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: too many streams");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_CVR, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_cvr_device<double>(*M, m, nnz, d_rst, d_col, (const double*)d_val, nstreams, s)
        : rvvla::build_cvr_device<float>(*M, m, nnz, d_rst, d_col, (const float*)d_val, nstreams, s);
    if (rc) throw Error(rc, rvvla::last_error());
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_vhcc_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                           const int32_t* d_col, const void* d_val, int64_t nparts, void* stream,
                           rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (nparts <= 0) nparts = rvvla::default_parts(m, nnz, rvvla::kVhccPartEntries);  // the convert default
    if (nparts > 2147483647LL) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: too many partitions");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_VHCC, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_vhcc_device<double>(*M, m, nnz, d_rst, d_col, (const double*)d_val, nparts, s)
        : rvvla::build_vhcc_device<float>(*M, m, nnz, d_rst, d_col, (const float*)d_val, nparts, s);
    if (rc) throw Error(rc, rvvla::last_error());
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_csr5_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                           const int32_t* d_col, const void* d_val, int sigma, void* stream,
                           rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    if (sigma <= 0) sigma = rvvla::default_csr5_sigma(m, nnz);
    if (sigma > 64) throw Error(RVVLA_ERR_CONVERSION, "CSR5: sigma must be in [1, 64]");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_CSR5, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_csr5_device<double>(*M, m, nnz, d_rst, d_col, (const double*)d_val, sigma, s)
        : rvvla::build_csr5_device<float>(*M, m, nnz, d_rst, d_col, (const float*)d_val, sigma, s);
    if (rc) throw Error(rc, rvvla::last_error());
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_hcsr_from_device(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* d_rst,
                           const int32_t* d_col, const void* d_val, int64_t R, int64_t C,
                           double p, void* stream, rvvla_mat_t* out) {
  return guarded([&]() {
    if (!out) throw Error(RVVLA_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: shape out of the int32 range");
    if (!d_rst || (nnz > 0 && (!d_col || !d_val)))
      throw Error(RVVLA_ERR_INVALID_ARG, "CSR: null device array");
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) throw Error(RVVLA_ERR_VALUE_TYPE, "unknown dtype");
    // reference defaults: R = C = 2000 (FP64), 4000 (FP32), p = 2
    // (reference src/scalar/crs_to_hcsr.cpp:141-143, :158-160)
    const int64_t dflt = dtype == RVVLA_F64 ? 2000 : 4000;
    if (R <= 0) R = dflt;
    if (C <= 0) C = dflt;
    if (!(p > 0.0)) p = 2.0;
    if (C > 65536) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: C must not exceed 65536 (16-bit col_idx)");
    const int64_t bm = (m + R - 1) / R, bn = (n + C - 1) / C;
    if (bn != 0 && bm > (int64_t(2147483646) / bn))
      throw Error(RVVLA_ERR_LENGTH, "HCSR: more than 2^31 - 2 blocks");
    if (rvvla::g_host_validation.load())
      throw Error(RVVLA_ERR_NO_DEVICE, "device-side conversion needs a device");
    auto M = rvvla::new_mat(RVVLA_FMT_HCSR, dtype, m, n, nnz);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = rvvla::check_csr_device(*M, m, n, nnz, d_rst, d_col, s)) throw Error(rc, rvvla::last_error());
    const int rc = dtype == RVVLA_F64
        ? rvvla::build_hcsr_device<double>(*M, m, n, nnz, d_rst, d_col, (const double*)d_val, R, C, p, s)
        : rvvla::build_hcsr_device<float>(*M, m, n, nnz, d_rst, d_col, (const float*)d_val, R, C, p, s);
    if (rc) throw Error(rc, rvvla::last_error());
    // the kernels' image: the layout copied back and laid out by the host
    // image builder
    if (dtype == RVVLA_F64) rvvla::hcsr_image_from_device<double>(*M, s);
    else rvvla::hcsr_image_from_device<float>(*M, s);
    *out = M.release();
    return RVVLA_OK;
  });
}

int rvvla_mat_read(rvvla_mat_t A, const char* field, void* dst, int64_t capacity, int64_t* count,
                   int* elem_bytes) {
  if (!A || !field || !count || !elem_bytes) return RVVLA_ERR_INVALID_ARG;
  const std::string f = field;
  const int vb = A->dtype == RVVLA_F64 ? 8 : A->dtype == RVVLA_F32 ? 4 : 2;
  struct V { const void* p; int64_t n; int e; };
  V v{nullptr, -1, 0};
  if (A->format == RVVLA_FMT_SELL) {
    const rvvla::DevSELL& d = A->sell;
    int64_t slots = 0;
    if (d.nslices > 0) {
      int32_t cs = 0, cl = 0, bs = 0;
      const int64_t l = d.nslices - 1;
      if (A->host_only) {
        cs = d.cs[l];
        cl = d.cl[l];
        bs = d.bs[l];
      } else {
        int rc = rvvla::hip_check(hipMemcpy(&cs, d.cs + l, 4, hipMemcpyDeviceToHost), "read");
        if (!rc) rc = rvvla::hip_check(hipMemcpy(&cl, d.cl + l, 4, hipMemcpyDeviceToHost), "read");
        if (!rc) rc = rvvla::hip_check(hipMemcpy(&bs, d.bs + l, 4, hipMemcpyDeviceToHost), "read");
        if (rc) return rc;
      }
      slots = (int64_t)cs + (int64_t)cl * bs;
    }
    v = f == "vCl" ? V{d.cl, d.nslices, 4} : f == "vCs" ? V{d.cs, d.nslices, 4}
      : f == "vBs" ? V{d.bs, d.nslices, 4} : f == "vSBs" ? V{d.sbs, d.nslices, 4}
      : f == "vCol" ? V{d.col, slots, 4} : f == "vPerm" ? V{d.perm, A->m, 4}
      : f == "vVal" ? V{d.val, slots, vb} : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_CSR) {
    const rvvla::DevCSR& c = A->csr;
    v = f == "rst" ? V{c.rst, A->m + 1, 4} : f == "col" ? V{c.col, A->nnz, 4}
      : f == "val" ? V{c.val, A->nnz, vb} : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_LAV && A->lav.device_built) {
    // built on the device: the layout of convert(RVVLA_FMT_LAV)
    const rvvla::DevLAV& c = A->lav;
    const int64_t S = c.nseg, cc = c.nchunks, tot = c.total_steps;
    v = f == "chunk_offsets" ? V{c.coff, S * (cc + 1), 4}
      : f == "chunk_offsets_full_mask" ? V{c.ref_cfull, S * cc, 4}
      : f == "segment_ptr" ? V{c.sptr, S + 1, 8}
      : f == "out_order" ? V{c.out_order, S * cc * 64, 4} : f == "mask" ? V{c.ref_mask, tot, 2}
      : f == "val" ? V{c.val, tot * 64, vb} : f == "col_id" ? V{c.col, tot * 64, 4}
      : f == "sparse_part.Rst" ? V{A->csr.rst, A->m + 1, 4}
      : V{nullptr, -1, 0};
    if (f == "sparse_part.Col" || f == "sparse_part.Val") {
      int32_t last = 0;
      int rc = A->host_only ? (last = A->csr.rst[A->m], 0)
                            : rvvla::hip_check(hipMemcpy(&last, A->csr.rst + A->m, 4, hipMemcpyDeviceToHost), "read");
      if (rc) return rc;
      v = f == "sparse_part.Col" ? V{A->csr.col, last, 4} : V{A->csr.val, last, vb};
    }
  } else if (A->format == RVVLA_FMT_VNEC) {
    const rvvla::DevVNEC& c = A->vnec;
    v = f == "cx" ? V{c.cx, c.nparts + 1, 4} : f == "cy" ? V{c.cy, c.nparts + 1, 4}
      : f == "nec_ptr" ? V{c.nec_ptr, c.nparts + 1, 4} : f == "use_x_indices" ? V{c.use_x, c.nlist, 4}
      : f == "ecr_indices" ? V{c.ecr, A->nnz, 2} : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_CVR) {
    const rvvla::DevCVR& c = A->cvr;
    const int64_t S = c.nstreams;
    v = f == "pack_start" ? V{c.pack_start, S + 1, 8} : f == "rec_start" ? V{c.rec_start, S, 4}
      : f == "flush" ? V{c.flush, c.npacks, 8} : f == "rec_rows" ? V{c.rec_rows, c.nrec, 4}
      : f == "tail_rows" ? V{c.tail_rows, S * 64, 4} : f == "col" ? V{c.col, c.npacks * 64, 4}
      : f == "val" ? V{c.val, c.npacks * 64, vb} : f == "empty_rows" ? V{c.empty_rows, c.nempty, 4}
      : f == "split_rows" ? V{c.split_rows, c.nsplit, 4} : f == "split_ptr" ? V{c.split_ptr, 2 * c.nsplit, 4}
      : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_VHCC) {
    const rvvla::DevVHCC& c = A->vhcc;
    const int64_t T = c.nparts, slots = c.nvec * 64;
    v = f == "start_vec" ? V{c.start_vec, T + 1, 4} : f == "rbase" ? V{c.rbase, T, 4}
      : f == "eor" ? V{c.eor, c.nvec, 8} : f == "row_arr" ? V{c.row_arr, c.nends, 4}
      : f == "fwd_row" ? V{c.fwd_row, T, 4} : f == "pass" ? V{c.pass, T, 1}
      : f == "col" ? V{c.col, slots, 4} : f == "val" ? V{c.val, slots, vb}
      : f == "empty_rows" ? V{c.empty_rows, c.nempty, 4} : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_CSR5) {
    const rvvla::DevCSR5& c = A->csr5;
    v = f == "_csr5_tile_pointer" ? V{c.tile_ptr, c.np + 1, 4}
      : f == "_csr5_tile_descriptor" ? V{c.desc, c.ntiles * 64 * c.dl, 4}
      : f == "_csr5_tile_empty_offset_pointer" ? V{c.eptr, c.np + 1, 4}
      : f == "_csr5_tile_empty_offset" ? V{c.eoff, c.noffsets, 4}
      : f == "_csr_column_index" ? V{c.col, A->nnz, 4} : f == "_csr_value" ? V{c.val, A->nnz, vb}
      : V{nullptr, -1, 0};
  } else if (A->format == RVVLA_FMT_HCSR && A->hcsr.ref_block_ptr) {
    // built on the device: the reference layout
    const rvvla::DevHCSR& h = A->hcsr;
    const int64_t nb = h.b_m * h.b_n + 1;
    v = f == "block_ptr" ? V{h.ref_block_ptr, nb, 4} : f == "row_offset" ? V{h.ref_row_offset, nb, 4}
      : f == "row_ptr" ? V{h.ref_row_ptr, h.nref_row_ptr, 4} : f == "col_idx" ? V{h.ref_col_idx, A->nnz, 2}
      : f == "values" ? V{h.ref_values, A->nnz, vb} : V{nullptr, -1, 0};
  }
  if (v.n < 0) {
    rvvla::set_last_error("rvvla_mat_read: unknown field " + f + " for this format");
    return RVVLA_ERR_NOT_FOUND;
  }
  *count = v.n;
  *elem_bytes = v.e;
  if (!dst) return RVVLA_OK;  // size query
  if (capacity < v.n * v.e) return RVVLA_ERR_LENGTH;
  if (v.n == 0) return RVVLA_OK;
  if (A->host_only) {
    std::memcpy(dst, v.p, (size_t)(v.n * v.e));
    return RVVLA_OK;
  }
  return rvvla::hip_check(hipMemcpy(dst, v.p, (size_t)(v.n * v.e), hipMemcpyDeviceToHost), "read");
}

int rvvla_set_host_validation(int on) {
  rvvla::g_host_validation.store(on ? 1 : 0);
  return RVVLA_OK;
}

int rvvla_mat_view(rvvla_mat_t A, const char* field, const void** data, int64_t* count,
                   int* elem_bytes) {
  if (!A || !field || !data || !count || !elem_bytes) return RVVLA_ERR_INVALID_ARG;
  if (!A->host_only) {
    rvvla::set_last_error("rvvla_mat_view needs a matrix uploaded in host-validation mode");
    return RVVLA_ERR_LOGIC;
  }
  if (A->format != RVVLA_FMT_CSR && A->format != RVVLA_FMT_HCSR) {
    rvvla::set_last_error("rvvla_mat_view: CSR and HCSR matrices only");
    return RVVLA_ERR_UNKNOWN_FORMAT;
  }
  const rvvla::DevCSR& c = A->csr;
  const rvvla::DevXSlice& x = c.xs;
  const rvvla::DevHCSR& h = A->hcsr;
  const int vb = A->dtype == RVVLA_F64 ? 8 : A->dtype == RVVLA_F32 ? 4 : 2;
  const int64_t nblk = 2 * c.nblocks;
  const int64_t xnnz = x.P ? A->nnz : 0, xpairs = x.P ? x.pairs + 1 : 0;
  const std::string f = field;
  struct V { const void* p; int64_t n; int e; };
  const bool hc = A->format == RVVLA_FMT_HCSR;
  const V v = hc ? (f == "hc.unit"    ? V{h.unit, h.nunits * rvvla::kHcUnitInts, 4}
                  : f == "hc.dxoff" ? V{h.dxoff, h.ndrows, 4}
                  : f == "hc.dcol"  ? V{h.dcol, h.ndrows * 64, 2}
                  : f == "hc.dval"  ? V{h.dval, h.ndrows * 64, vb}
                  : f == "hc.tcol"  ? V{h.tcol, h.ntail, 4}
                  : f == "hc.trow"  ? V{h.trow, h.ntail, 1}
                  : f == "hc.tval"  ? V{h.tval, h.ntail, vb}
                  : f == "hc.split" ? V{h.split, h.nsplit * 4, 4}
                                    : V{nullptr, -1, 0})
            : f == "rst"     ? V{c.rst, A->m + 1, 4}
            : f == "col"     ? V{c.col, A->nnz, 4}
            : f == "blocks"  ? V{c.blocks, nblk, 4}
            : f == "xs.rst"  ? V{x.rst, xpairs, 4}
            : f == "xs.col"  ? V{x.col, xnnz, 4}
            : f == "xs.val"  ? V{x.val, xnnz, vb}
            : f == "xs.blk"  ? V{x.blk, 2 * x.nblk, 4}
            : f == "xs.mask" ? V{x.mask, x.P ? A->m : 0, 2}
            : f == "xs.base" ? V{x.base, x.P ? x.P + 1 : 0, 4}
            : f == "xs.pref" ? V{x.pref, x.P * x.nq, 4}
            : f == "xl.col"    ? V{c.xl.col, c.xl.S ? c.xl.nent : 0, 2}
            : f == "xl.val"    ? V{c.xl.val, c.xl.S ? c.xl.nent : 0, vb}
            : f == "xl.task"   ? V{c.xl.task, 4 * c.xl.ntasks, 4}
            : f == "xl.units"  ? V{c.xl.units, 3 * c.xl.nunits, 4}
            : f == "xl.wg"     ? V{c.xl.wg, c.xl.S ? c.xl.nwg + 1 : 0, 4}
            : f == "xl.rowptr" ? V{c.xl.rowptr, c.xl.S ? A->m + 1 : 0, 4}
            : f == "xl.cblk"   ? V{c.xl.cblk, 2 * c.xl.nblocks, 4}
            : f == "xl.off"    ? V{c.xl.off, c.xl.S * (c.xl.nblocks + 1), 4}
            : f == "xl.lpos"   ? V{c.xl.lpos, c.xl.npieces, 2}
            : f == "sell.cl"   ? V{A->sell.cl, c.sell ? A->sell.nslices : 0, 4}
            : f == "sell.vPerm" ? V{A->sell.perm, c.sell && A->sell.perm ? A->m : 0, 4}
                               : V{nullptr, -1, 0};
  if (v.n < 0) {
    rvvla::set_last_error("rvvla_mat_view: unknown field " + f);
    return RVVLA_ERR_NOT_FOUND;
  }
  *data = v.p;
  *count = v.n;
  *elem_bytes = v.e;
  return RVVLA_OK;
}

int rvvla_mat_free(rvvla_mat_t A) {
  delete A;
  return RVVLA_OK;
}

int rvvla_mat_info(rvvla_mat_t A, int64_t* m, int64_t* n, int64_t* nnz, int* format, int* dtype,
                   int64_t* device_bytes) {
  if (!A) return RVVLA_ERR_INVALID_ARG;
  if (m) *m = A->m;
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  if (format) *format = A->format;
  if (dtype) *dtype = A->dtype;
  if (device_bytes) *device_bytes = A->device_bytes;
  return RVVLA_OK;
}

static int need_device(rvvla_mat_t A) {
  if (!A) return RVVLA_ERR_INVALID_ARG;
  if (A->host_only) {
    rvvla::set_last_error("matrix was uploaded in host-validation mode");
    return RVVLA_ERR_NO_DEVICE;
  }
  return RVVLA_OK;
}

extern "C++" {
namespace rvvla {

int64_t spmv_scratch_bytes(const rvvla_mat_s& A) {
  const int64_t acc = A.dtype == RVVLA_F64 ? 8 : 4;
  auto sell_part = [&](const DevSELL& d) { return d.nsplit > 0 ? d.npieces * 64 * acc : 0; };
  switch (A.format) {
    case RVVLA_FMT_CSR:
      if (A.csr.sell) return sell_part(A.sell);
      if (A.csr.xl.S > 0) return std::max<int64_t>(A.csr.xl.npieces, 1) * acc;
      if (A.csr.xs.P > 0) return A.csr.xs.pairs * acc;
      return 0;
    case RVVLA_FMT_SELL: return sell_part(A.sell);
    case RVVLA_FMT_HCSR: return A.hcsr.nparts * 64 * acc;
    case RVVLA_FMT_CVR: return A.cvr.nslots * acc;
    case RVVLA_FMT_LAV: return A.lav.nsplit > 0 ? A.lav.npieces * 64 * acc : 0;
    case RVVLA_FMT_CSR5: return A.csr5.ntiles * (2 * acc + 4) + 16;
    case RVVLA_FMT_VHCC: return A.vhcc.nparts * 2 * acc + 16;
    case RVVLA_FMT_VNEC: return A.vnec.nparts * 2 * acc + 16;
  }
  return 0;
}

int spmv_dispatch(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                  hipStream_t s, void* scratch) {
  switch (A.format) {
    case RVVLA_FMT_CSR:
      return A.csr.sell ? launch_sell(A, alpha, x, beta, y, s, scratch)
                        : launch_csr(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_HCSR: return launch_hcsr(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_SELL: return launch_sell(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_CSR5: return launch_csr5(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_CVR: return launch_cvr(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_VHCC: return launch_vhcc(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_LAV: return launch_lav(A, alpha, x, beta, y, s, scratch);
    case RVVLA_FMT_VNEC: return launch_vnec(A, alpha, x, beta, y, s, scratch);
  }
  return RVVLA_ERR_UNKNOWN_FORMAT;
}

int spmv_check_args(rvvla_mat_t A, const void* x, void* y) {
  if (!A || (A->n > 0 && !x) || (A->m > 0 && !y)) return RVVLA_ERR_INVALID_ARG;
  return need_device(A);
}

}  // namespace rvvla
}  // extern "C++"

int rvvla_spmv(rvvla_mat_t A, double alpha, const void* x, double beta, void* y, void* stream) {
  if (int rc = rvvla::spmv_check_args(A, x, y)) return rc;
  return rvvla::spmv_dispatch(*A, alpha, x, beta, y, (hipStream_t)stream, nullptr);
}

// CSR image of the transpose from CSR arrays on the host: counting sort of
// the entries by column (a column's entries stay in row order), values
// moved with them as raw storage elements (es bytes each), adaptive blocks
// over the n rows of the transpose.
static void put_transposed_image(rvvla_mat_s& A, const int32_t* rst, const int32_t* col,
                                 const unsigned char* val, size_t es) {
  const int64_t m = A.m, n = A.n, nnz = A.nnz;
  std::vector<int32_t> ptr((size_t)n + 1, 0), row((size_t)nnz);
  std::vector<unsigned char> tval((size_t)nnz * es);
  for (int64_t e = 0; e < nnz; ++e) ptr[(size_t)col[e] + 1]++;
  for (int64_t c = 0; c < n; ++c) ptr[(size_t)c + 1] += ptr[(size_t)c];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t r = 0; r < m; ++r)
    for (int32_t e = rst[r]; e < rst[r + 1]; ++e) {
      const int32_t p = fill[(size_t)col[e]]++;
      row[(size_t)p] = (int32_t)r;
      std::memcpy(&tval[(size_t)p * es], &val[(size_t)e * es], es);
    }
  const std::vector<int32_t> blocks = rvvla::csr_row_blocks(ptr.data(), 0, n);
  rvvla::DevCSR& t = A.tcsr;
  t.rst = A.put_vec(ptr);
  t.col = A.put_vec(row);
  t.val = A.put(tval.data(), tval.size());
  t.blocks = A.put_vec(blocks);
  t.nblocks = (int64_t)blocks.size() / 2;
  A.has_tcsr = true;
}

extern "C++" {
namespace rvvla {
template <class V>
void attach_transpose_source(rvvla_mat_s& M, std::shared_ptr<const spMtxCRS<V>> src) {
  M.transpose_source = [src](rvvla_mat_s& A) {
    const spMtxCRS<V>& c = *src;
    if ((int64_t)c.m != A.m || (int64_t)c.n != A.n || (int64_t)c.nz != A.nnz)
      throw Error(RVVLA_ERR_LOGIC, "transpose: the source CSR does not match the matrix");
    // values in the matrix's storage type, rounded as its upload rounds them
    const size_t es = A.dtype == RVVLA_F64 ? 8 : A.dtype == RVVLA_F32 ? 4 : 2;
    std::vector<unsigned char> v((size_t)c.nz * es);
    for (size_t e = 0; e < c.nz; ++e) {
      if (A.dtype == RVVLA_F64) {
        const double d = (double)c.Val[e];
        std::memcpy(&v[e * 8], &d, 8);
      } else if (A.dtype == RVVLA_F32) {
        const float f = (float)c.Val[e];
        std::memcpy(&v[e * 4], &f, 4);
      } else {
        const uint16_t h = A.dtype == RVVLA_BF16 ? f32_to_bf16((float)c.Val[e]) : f32_to_f16((float)c.Val[e]);
        std::memcpy(&v[e * 2], &h, 2);
      }
    }
    put_transposed_image(A, c.Rst, c.Col, v.data(), es);
  };
}
template void attach_transpose_source<double>(rvvla_mat_s&, std::shared_ptr<const spMtxCRS<double>>);
template void attach_transpose_source<float>(rvvla_mat_s&, std::shared_ptr<const spMtxCRS<float>>);
}  // namespace rvvla
}  // extern "C++"

// The CSR image of A^T: a CSR matrix's own arrays copied back, or the
// source CSR a matrix of another format was converted from.
static int build_transposed_image(rvvla_mat_t A) {
  if (A->has_tcsr) return RVVLA_OK;
  return guarded([&]() {
    if (A->format != RVVLA_FMT_CSR) {
      if (!A->transpose_source)
        throw Error(RVVLA_ERR_UNSUPPORTED_SPMV,
                    "the transposed product needs the CSR this matrix was converted from: "
                    "matrices converted from CSR (rvvla_convert, the C++ convert) keep it; "
                    "matrices uploaded from format arrays or built on the device do not");
      A->transpose_source(*A);
      return RVVLA_OK;
    }
    const int64_t m = A->m, nnz = A->nnz;
    const size_t es = A->dtype == RVVLA_F64 ? 8 : A->dtype == RVVLA_F32 ? 4 : 2;
    std::vector<int32_t> rst((size_t)m + 1), col((size_t)nnz);
    std::vector<unsigned char> val((size_t)nnz * es);
    int rc = rvvla::hip_check(hipMemcpy(rst.data(), A->csr.rst, rst.size() * 4, hipMemcpyDeviceToHost), "copy Rst");
    if (!rc && nnz) rc = rvvla::hip_check(hipMemcpy(col.data(), A->csr.col, col.size() * 4, hipMemcpyDeviceToHost), "copy Col");
    if (!rc && nnz) rc = rvvla::hip_check(hipMemcpy(val.data(), A->csr.val, val.size(), hipMemcpyDeviceToHost), "copy Val");
    if (rc) throw Error(rc, rvvla::last_error());
    put_transposed_image(*A, rst.data(), col.data(), val.data(), es);
    return RVVLA_OK;
  });
}

int rvvla_spmv_op(rvvla_mat_t A, int op, double alpha, const void* x, double beta, void* y,
                  void* stream) {
  if (op == 0) return rvvla_spmv(A, alpha, x, beta, y, stream);
  if (op != 1 || !A) return RVVLA_ERR_INVALID_ARG;
  if (int rc = need_device(A)) return rc;
  if ((A->m > 0 && !x) || (A->n > 0 && !y)) return RVVLA_ERR_INVALID_ARG;
  if (int rc = build_transposed_image(A)) return rc;
  return rvvla::launch_csr_image(A->tcsr, A->n, A->dtype, alpha, x, beta, y, (hipStream_t)stream);
}

int rvvla_spmm(rvvla_mat_t A, int k, double alpha, const void* X, int64_t ldx, double beta,
               void* Y, int64_t ldy, void* stream) {
  if (int rc = need_device(A)) return rc;
  if (!A || k < 0 || ldx < k || ldy < k || (k > 0 && ((A->n > 0 && !X) || (A->m > 0 && !Y))))
    return RVVLA_ERR_INVALID_ARG;
  if (A->format != RVVLA_FMT_CSR) {
    rvvla::set_last_error("SpMM is implemented for the CSR format");
    return RVVLA_ERR_UNSUPPORTED_SPMV;
  }
  return rvvla::launch_csr_spmm(*A, k, alpha, X, ldx, beta, Y, ldy, (hipStream_t)stream);
}

static int need_csr_real(rvvla_mat_t A) {
  if (int rc = need_device(A)) return rc;
  if (A->format != RVVLA_FMT_CSR) {
    rvvla::set_last_error("differentiable SpMM needs a CSR matrix");
    return RVVLA_ERR_UNSUPPORTED_SPMV;
  }
  if (A->dtype != RVVLA_F64 && A->dtype != RVVLA_F32) {
    rvvla::set_last_error("differentiable SpMM needs FP32 or FP64 values");
    return RVVLA_ERR_VALUE_TYPE;
  }
  return RVVLA_OK;
}

int rvvla_prepare_transpose(rvvla_mat_t A) {
  if (int rc = need_csr_real(A)) return rc;
  if (A->csc.ptr) return RVVLA_OK;
  return guarded([&]() {
    const int64_t m = A->m, n = A->n, nnz = A->nnz;
    std::vector<int32_t> rst((size_t)m + 1), col((size_t)nnz);
    int rc = rvvla::hip_check(hipMemcpy(rst.data(), A->csr.rst, rst.size() * 4, hipMemcpyDeviceToHost), "copy Rst");
    if (!rc && nnz)
      rc = rvvla::hip_check(hipMemcpy(col.data(), A->csr.col, col.size() * 4, hipMemcpyDeviceToHost), "copy Col");
    if (rc) throw Error(rc, rvvla::last_error());
    // counting sort by column; entries of a column stay in row order
    std::vector<int32_t> ptr((size_t)n + 1, 0), row((size_t)nnz), perm((size_t)nnz);
    for (int64_t e = 0; e < nnz; ++e) ptr[(size_t)col[(size_t)e] + 1]++;
    for (int64_t c = 0; c < n; ++c) ptr[(size_t)c + 1] += ptr[(size_t)c];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t r = 0; r < m; ++r)
      for (int32_t e = rst[(size_t)r]; e < rst[(size_t)r + 1]; ++e) {
        const int32_t p = fill[(size_t)col[(size_t)e]]++;
        row[(size_t)p] = (int32_t)r;
        perm[(size_t)p] = e;
      }
    A->csc.ptr = A->put_vec(ptr);
    A->csc.row = A->put_vec(row);
    A->csc.perm = A->put_vec(perm);
    return RVVLA_OK;
  });
}

int rvvla_spmm_values(rvvla_mat_t A, const void* vals, int k, double alpha, const void* X,
                      int64_t ldx, double beta, void* Y, int64_t ldy, void* stream) {
  if (int rc = need_csr_real(A)) return rc;
  if (k < 0 || ldx < k || ldy < k || (k > 0 && ((A->n > 0 && !X) || (A->m > 0 && !Y))))
    return RVVLA_ERR_INVALID_ARG;
  return rvvla::launch_spmm_values(*A, false, vals ? vals : A->csr.val, k, alpha, X, ldx, beta, Y,
                                   ldy, (hipStream_t)stream);
}

int rvvla_spmm_transpose(rvvla_mat_t A, const void* vals, int k, double alpha, const void* X,
                         int64_t ldx, double beta, void* Y, int64_t ldy, void* stream) {
  if (int rc = need_csr_real(A)) return rc;
  if (k < 0 || ldx < k || ldy < k || (k > 0 && ((A->m > 0 && !X) || (A->n > 0 && !Y))))
    return RVVLA_ERR_INVALID_ARG;
  if (int rc = rvvla_prepare_transpose(A)) return rc;
  return rvvla::launch_spmm_values(*A, true, vals ? vals : A->csr.val, k, alpha, X, ldx, beta, Y,
                                   ldy, (hipStream_t)stream);
}

int rvvla_spmm_backward(rvvla_mat_t A, const void* vals, int k, const void* DY, int64_t lddy,
                        const void* X, int64_t ldx, void* dX, int64_t lddx, void* dvals,
                        void* stream) {
  if (int rc = need_csr_real(A)) return rc;
  if (k < 0 || lddy < k || ldx < k || (dX && lddx < k) ||
      (k > 0 && A->nnz > 0 && (!DY || (dvals && !X))) || (k > 0 && A->n > 0 && dX && !DY))
    return RVVLA_ERR_INVALID_ARG;
  if (!dX && !dvals) return RVVLA_OK;
  if (k == 0) {  // no columns: dX is n x 0, and every <DY[r, :], X[c, :]> is 0
    if (dvals && A->nnz > 0)
      return rvvla::hip_check(hipMemsetAsync(dvals, 0, (size_t)A->nnz * (A->dtype == RVVLA_F64 ? 8 : 4),
                                             (hipStream_t)stream), "clear dvals");
    return RVVLA_OK;
  }
  if (int rc = rvvla_prepare_transpose(A)) return rc;
  return rvvla::launch_spmm_backward(*A, vals ? vals : A->csr.val, k, DY, lddy, X, ldx, dX, lddx,
                                     dvals, (hipStream_t)stream);
}

int rvvla_sddmm(rvvla_mat_t A, int k, double alpha, const void* DY, int64_t lddy, const void* X,
                int64_t ldx, void* out_vals, void* stream) {
  if (int rc = need_csr_real(A)) return rc;
  if (k < 0 || lddy < k || ldx < k || (A->nnz > 0 && (!DY || !X || !out_vals)))
    return RVVLA_ERR_INVALID_ARG;
  return rvvla::launch_sddmm(*A, k, DY, lddy, X, ldx, alpha, out_vals, (hipStream_t)stream);
}

// Staging buffers hold max(m, n) elements each, so that either operation
// can use them: op 0 reads n entries of b and writes m of y, op 1 the
// reverse.
int rvvla_stage_in_op(rvvla_mat_t A, int op, const void* b_host, const void* y_host) {
  if (int rc = need_device(A)) return rc;
  if (op != 0 && op != 1) return RVVLA_ERR_INVALID_ARG;
  const int64_t nb = op ? A->m : A->n, ny = op ? A->n : A->m;
  if (nb > 0 && !b_host) return RVVLA_ERR_INVALID_ARG;
  const size_t es = A->dtype == RVVLA_F64 ? 8 : 4;
  const size_t cap = (size_t)std::max(A->m, A->n) * es + 8;
  int rc;
  if (!A->xstage && (rc = rvvla::hip_check(hipMalloc(&A->xstage, cap), "hipMalloc x"))) return rc;
  if (!A->ystage && (rc = rvvla::hip_check(hipMalloc(&A->ystage, cap), "hipMalloc y"))) return rc;
  if ((rc = rvvla::hip_check(hipMemcpy(A->xstage, b_host, (size_t)nb * es, hipMemcpyHostToDevice), "copy b")))
    return rc;
  if (y_host)
    return rvvla::hip_check(hipMemcpy(A->ystage, y_host, (size_t)ny * es, hipMemcpyHostToDevice), "copy y");
  return rvvla::hip_check(hipMemset(A->ystage, 0, (size_t)ny * es), "clear y");
}

int rvvla_mult_staged_op(rvvla_mat_t A, int op, double alpha, double beta) {
  if (!A) return RVVLA_ERR_INVALID_ARG;
  if (!A->xstage || !A->ystage) {
    rvvla::set_last_error("MULT before PREPARATION");
    return RVVLA_ERR_LOGIC;
  }
  if (int rc = rvvla_spmv_op(A, op, alpha, A->xstage, beta, A->ystage, nullptr)) return rc;
  return rvvla::hip_check(hipDeviceSynchronize(), "multiply");
}

int rvvla_stage_out_op(rvvla_mat_t A, int op, void* y_host) {
  if (!A || (op != 0 && op != 1)) return RVVLA_ERR_INVALID_ARG;
  const int64_t ny = op ? A->n : A->m;
  if (ny > 0 && !y_host) return RVVLA_ERR_INVALID_ARG;
  if (!A->ystage) {
    rvvla::set_last_error("GET_RESULTS before PREPARATION");
    return RVVLA_ERR_LOGIC;
  }
  const size_t es = A->dtype == RVVLA_F64 ? 8 : 4;
  return rvvla::hip_check(hipMemcpy(y_host, A->ystage, (size_t)ny * es, hipMemcpyDeviceToHost), "copy y back");
}

int rvvla_spmv_host_op(rvvla_mat_t A, int op, double alpha, const void* b_host, double beta,
                       void* y_host) {
  if (int rc = rvvla_stage_in_op(A, op, b_host, y_host)) return rc;
  if (int rc = rvvla_mult_staged_op(A, op, alpha, beta)) return rc;
  return rvvla_stage_out_op(A, op, y_host);
}

int rvvla_stage_in(rvvla_mat_t A, const void* b_host, const void* y_host) {
  return rvvla_stage_in_op(A, 0, b_host, y_host);
}
int rvvla_mult_staged(rvvla_mat_t A, double alpha, double beta) {
  return rvvla_mult_staged_op(A, 0, alpha, beta);
}
int rvvla_stage_out(rvvla_mat_t A, void* y_host) { return rvvla_stage_out_op(A, 0, y_host); }
int rvvla_spmv_host(rvvla_mat_t A, double alpha, const void* b_host, double beta, void* y_host) {
  if (!A || (A->n > 0 && !b_host) || (A->m > 0 && !y_host)) return RVVLA_ERR_INVALID_ARG;
  return rvvla_spmv_host_op(A, 0, alpha, b_host, beta, y_host);
}

}  // extern "C"
