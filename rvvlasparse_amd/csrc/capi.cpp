// capi.cpp -- the extern "C" boundary of librvvla_hip.so.
//
// Conversion builds the reference host containers (include/*.h), upload
// validates every index against the matrix shape on the host (so no kernel
// can be handed an out-of-range gather) and copies the device image once,
// and rvvla_spmv dispatches to the format's gfx950 kernel on the caller's
// stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
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

}  // namespace rvvla

using rvvla::Error;

// ---------------------------------------------------------------- device matrix
rvvla_mat_s::~rvvla_mat_s() {
  for (void* p : allocs) (void)hipFree(p);
  if (xstage) (void)hipFree(xstage);
  if (ystage) (void)hipFree(ystage);
}

void* rvvla_mat_s::put(const void* host, size_t bytes) {
  void* d = nullptr;
  const size_t nb = bytes ? bytes : 4;
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
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) throw Error(RVVLA_ERR_NO_DEVICE, last_error());
  auto M = std::make_unique<rvvla_mat_s>();
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

// Adaptive row blocks (see k_csr.hip).
static std::vector<int32_t> csr_blocks(int64_t m, const int* rst) {
  std::vector<int32_t> b;
  b.reserve((size_t)(m / 64 + 2));
  b.push_back(0);
  int64_t r = 0;
  while (r < m) {
    if (rst[r + 1] - rst[r] > kLongRow) {
      b.push_back((int32_t)++r);
      continue;
    }
    const int64_t start = r;
    int64_t cnt = 0;
    while (r < m && r - start < kStreamRows) {
      const int64_t len = rst[r + 1] - rst[r];
      if (len > kLongRow || cnt + len > kStreamCap) break;
      cnt += len;
      ++r;
    }
    b.push_back((int32_t)r);
  }
  return b;
}

template <class V>
rvvla_mat_t upload_csr(int dtype, const spMtxCRS<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  check_csr((int64_t)a.m, (int64_t)a.n, (int64_t)a.nz, a.Rst, a.Col);
  auto M = new_mat(RVVLA_FMT_CSR, dtype, (int64_t)a.m, (int64_t)a.n, (int64_t)a.nz);
  const std::vector<int32_t> blocks = csr_blocks((int64_t)a.m, a.Rst);
  M->csr.rst = (const int32_t*)M->put(a.Rst, (a.m + 1) * sizeof(int));
  M->csr.col = (const int32_t*)M->put(a.Col, a.nz * sizeof(int));
  M->csr.val = put_values(*M, dtype, a.Val, a.nz);
  M->csr.blocks = M->put_vec(blocks);
  M->csr.nblocks = (int64_t)blocks.size() - 1;
  return M.release();
}

template <class V>
rvvla_mat_t upload_hcsr(int dtype, const spMtxHCSR<V>& a) {
  check_dtype_for(dtype, std::is_same<V, double>::value);
  const int64_t m = (int64_t)a.m, n = (int64_t)a.n, nz = (int64_t)a.nz;
  const int64_t R = (int64_t)a.R, C = (int64_t)a.C;
  if (R <= 0 || C <= 0 || C > 65536) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: bad R or C");
  const size_t acc_bytes = dtype == RVVLA_F64 ? 8 : 4;
  if ((size_t)R * acc_bytes > 65536)
    throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "HCSR: R * accumulator size must fit 64 KiB of LDS");
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
  std::vector<int32_t> brow_ptr((size_t)bm + 1, 0), nb_col, nb_ptr;
  std::vector<uint32_t> nb_roff, rp32;
  std::vector<uint16_t> coo16((size_t)nz, 0);
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
        continue;
      }
      nb_col.push_back((int32_t)bc);
      nb_ptr.push_back((int32_t)e0);
      if (kind == 1) {
        if ((size_t)off + (size_t)rows + 1 > a.row_ptr.size())
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: row_ptr too short");
        const uint32_t* rp = a.row_ptr.data() + off;
        if (rp[0] != 0 || (int64_t)rp[rows] != e1 - e0)
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: local row_ptr does not span the block");
        for (int64_t r = 0; r < rows; ++r)
          if (rp[r + 1] < rp[r]) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: local row_ptr");
        nb_roff.push_back((1u << 30) | (uint32_t)rp32.size());
        rp32.insert(rp32.end(), rp, rp + rows + 1);
        if (rp32.size() > 0x3FFFFFFFu) throw Error(RVVLA_ERR_LENGTH, "HCSR: rp32 overflow");
      } else if (kind == 2) {
        if ((size_t)off + (size_t)(e1 - e0) > a.row_ptr.size())
          throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: row_ptr too short");
        for (int64_t e = e0; e < e1; ++e) {
          const uint32_t lr = a.row_ptr[off + (size_t)(e - e0)];
          if ((int64_t)lr >= rows) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: COO row out of block");
          coo16[(size_t)e] = (uint16_t)lr;
        }
        nb_roff.push_back(2u << 30);
      } else {
        throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: unknown block kind");
      }
    }
    brow_ptr[(size_t)br + 1] = (int32_t)nb_col.size();
  }
  nb_ptr.push_back((int32_t)nz);
  auto M = new_mat(RVVLA_FMT_HCSR, dtype, m, n, nz);
  DevHCSR& d = M->hcsr;
  d.R = R;
  d.C = C;
  d.b_m = bm;
  d.b_n = bn;
  d.brow_ptr = M->put_vec(brow_ptr);
  d.nb_col = M->put_vec(nb_col);
  d.nb_ptr = M->put_vec(nb_ptr);
  d.nb_roff = M->put_vec(nb_roff);
  d.rp32 = M->put_vec(rp32);
  d.coo16 = M->put_vec(coo16);
  d.col16 = M->put_vec(a.col_idx);
  d.val = put_values(*M, dtype, a.values.data(), a.values.size());
  return M.release();
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
  DevSELL& d = M->sell;
  d.nslices = nb;
  d.cl = M->put_vec(a.vCl);
  d.cs = M->put_vec(a.vCs);
  d.bs = M->put_vec(a.vBs);
  d.sbs = M->put_vec(a.vSBs);
  d.col = M->put_vec(a.vCol);
  d.perm = M->put_vec(a.vPerm);
  d.val = put_values(*M, dtype, a.vVal.data(), a.vVal.size());
  return M.release();
}

#define RVVLA_INST_UPLOAD(F, T) template rvvla_mat_t F<double>(int, const T<double>&); \
  template rvvla_mat_t F<float>(int, const T<float>&);
RVVLA_INST_UPLOAD(upload_csr, spMtxCRS)
RVVLA_INST_UPLOAD(upload_hcsr, spMtxHCSR)
RVVLA_INST_UPLOAD(upload_sell, spMtxSELL_C_Sigma)

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
      L.upload = [a, dtype]() { return upload_csr<V>(dtype, *a); };
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
                    param_int(P, "sigma", default_sell_sigma((int64_t)crs->m, (int64_t)crs->nz)));
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

int rvvla_spmv(rvvla_mat_t A, double alpha, const void* x, double beta, void* y, void* stream) {
  if (!A || (A->n > 0 && !x) || (A->m > 0 && !y)) return RVVLA_ERR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  switch (A->format) {
    case RVVLA_FMT_CSR: return rvvla::launch_csr(*A, alpha, x, beta, y, s);
    case RVVLA_FMT_HCSR: return rvvla::launch_hcsr(*A, alpha, x, beta, y, s);
    case RVVLA_FMT_SELL: return rvvla::launch_sell(*A, alpha, x, beta, y, s);
  }
  return RVVLA_ERR_UNKNOWN_FORMAT;
}

int rvvla_spmm(rvvla_mat_t A, int k, double alpha, const void* X, int64_t ldx, double beta,
               void* Y, int64_t ldy, void* stream) {
  if (!A || k < 0 || ldx < k || ldy < k || (k > 0 && ((A->n > 0 && !X) || (A->m > 0 && !Y))))
    return RVVLA_ERR_INVALID_ARG;
  if (A->format != RVVLA_FMT_CSR) {
    rvvla::set_last_error("SpMM is implemented for the CSR format");
    return RVVLA_ERR_UNSUPPORTED_SPMV;
  }
  return rvvla::launch_csr_spmm(*A, k, alpha, X, ldx, beta, Y, ldy, (hipStream_t)stream);
}

int rvvla_spmv_host(rvvla_mat_t A, double alpha, const void* b_host, double beta, void* y_host) {
  if (!A || (A->n > 0 && !b_host) || (A->m > 0 && !y_host)) return RVVLA_ERR_INVALID_ARG;
  const size_t es = A->dtype == RVVLA_F64 ? 8 : 4;
  int rc;
  if (!A->xstage && (rc = rvvla::hip_check(hipMalloc(&A->xstage, (size_t)A->n * es + 8), "hipMalloc x")))
    return rc;
  if (!A->ystage && (rc = rvvla::hip_check(hipMalloc(&A->ystage, (size_t)A->m * es + 8), "hipMalloc y")))
    return rc;
  if ((rc = rvvla::hip_check(hipMemcpy(A->xstage, b_host, (size_t)A->n * es, hipMemcpyHostToDevice), "copy b")))
    return rc;
  if ((rc = rvvla::hip_check(hipMemcpy(A->ystage, y_host, (size_t)A->m * es, hipMemcpyHostToDevice), "copy y")))
    return rc;
  if ((rc = rvvla_spmv(A, alpha, A->xstage, beta, A->ystage, nullptr))) return rc;
  return rvvla::hip_check(hipMemcpy(y_host, A->ystage, (size_t)A->m * es, hipMemcpyDeviceToHost), "copy y back");
}

}  // extern "C"
