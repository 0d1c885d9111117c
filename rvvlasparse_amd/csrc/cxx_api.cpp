// cxx_api.cpp -- the C++ drop-in: convert<T, Fmt, true> and
// sparse_mv<T, Fmt, true, stage> for the eight formats, on top of the C-ABI.
//
// Mirrors the reference API (reference include/spmv_mtx.h:21-29,
// include/convert.h:30-31): the same template names, arguments and
// std::vector signature.  convert builds the format's host fields and
// uploads its device image once (dist.device), or with "device" = 1 builds
// the image on the device and reads the host fields back from it;
// sparse_mv uploads lazily when a container was filled some other way.  Stages (reference
// include/types.h:33-38): PREPARATION copies b and y to the device,
// MULT runs the kernel, GET_RESULTS copies y back, ALL_STAGES does all
// three.  Status codes are the RVVLA_* codes (0 = success).  The reference
// ignores type_op; here SPARSE_OPERATION_TRANSPOSE computes A^T b: for CRS
// from the matrix itself, for the other formats from the copy of the CRS
// their convert keeps (a container filled some other way answers it with
// RVVLA_ERR_UNSUPPORTED_SPMV rather than computing the untransposed
// product).
#include <omp.h>

#include <algorithm>
#include <type_traits>

#include "devmat.h"
#include "internal.h"

namespace SparseMatrixLib {

namespace {

using rvvla::Error;

template <class F>
int guard(F&& f) {
  try {
    f();
    return RVVLA_OK;
  } catch (const Error& e) {
    rvvla::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    return RVVLA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    rvvla::set_last_error(e.what());
    return RVVLA_ERR_LOGIC;
  }
}

template <class T>
constexpr int dtype_of() {
  return std::is_same<T, double>::value ? RVVLA_F64 : RVVLA_F32;
}

// A converted matrix of a format other than CRS keeps a copy of its source
// CRS for the transposed product (its CSR image of A^T is built from it on
// first use).
template <class T>
void keep_source(const DeviceImage& d, const spMtxCRS<T>& s) {
  if (d) rvvla::attach_transpose_source<T>(*d.get(), std::make_shared<const spMtxCRS<T>>(s));
}

// Device image of each container (uploaded on first use).
template <class T>
rvvla_mat_t image(const spMtxCRS<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_csr<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxHCSR<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_hcsr<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxSELL_C_Sigma<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_sell<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxCSR5<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_csr5<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxLAV<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_lav<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxCVR<T>& a) {
  if (!a.device) {
    int64_t nz = 0;
    for (const T& v : a.vPack_vec_vals) nz += v != T(0);
    a.device.reset(rvvla::upload_cvr<T>(dtype_of<T>(), a, nz));
  }
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxVHCC<T>& a) {
  if (!a.device) a.device.reset(rvvla::upload_vhcc<T>(dtype_of<T>(), a));
  return a.device.get();
}
template <class T>
rvvla_mat_t image(const spMtxVNEC<T>& a) {
  if (!a.device) {
    if (!a.G) throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: no source matrix");
    a.device.reset(rvvla::upload_vnec<T>(dtype_of<T>(), a, *a.G));
  }
  return a.device.get();
}

template <class T, class M>
sparse_matrix_status run_stage(sparse_matrix_mv_stage st, sparse_operation_t op, T alpha, const M& mat,
                               const std::vector<T>& b, T beta, std::vector<T>& y) {
  sparse_matrix_status status;
  if (op != SPARSE_OPERATION_NON_TRANSPOSE && op != SPARSE_OPERATION_TRANSPOSE) {
    rvvla::set_last_error("only SPARSE_OPERATION_NON_TRANSPOSE and SPARSE_OPERATION_TRANSPOSE");
    status.code = RVVLA_ERR_UNSUPPORTED_SPMV;
    return status;
  }
  const int o = op == SPARSE_OPERATION_TRANSPOSE ? 1 : 0;
  status.code = guard([&]() {
    rvvla_mat_t A = image(mat);
    int64_t m = 0, n = 0;
    rvvla_mat_info(A, &m, &n, nullptr, nullptr, nullptr, nullptr);
    if ((int64_t)b.size() < (o ? m : n) || (int64_t)y.size() < (o ? n : m))
      throw Error(RVVLA_ERR_INVALID_ARG, "b or y shorter than the matrix");
    int rc = RVVLA_OK;
    switch (st) {
      case SPARSE_MATRIX_MV_PREPARATION: rc = rvvla_stage_in_op(A, o, b.data(), y.data()); break;
      case SPARSE_MATRIX_MV_MULT: rc = rvvla_mult_staged_op(A, o, alpha, beta); break;
      case SPARSE_MATRIX_MV_GET_RESULTS: rc = rvvla_stage_out_op(A, o, y.data()); break;
      case SPARSE_MATRIX_MV_ALL_STAGES:
        rc = rvvla_spmv_host_op(A, o, alpha, b.data(), beta, y.data());
        break;
    }
    if (rc) throw Error(rc, rvvla_strerror(rc));
  });
  return status;
}

}  // namespace

// ---------------------------------------------------------------- sparse_mv
#define RVVLA_SPMV(T, FMT, STAGE)                                                              \
  template <>                                                                                  \
  sparse_matrix_status sparse_mv<T, FMT, true, STAGE>(                                          \
      sparse_operation_t op, T alpha, const FMT<T>& mat, sparse_matrix_descr,                   \
      const std::vector<T>& b, T beta, std::vector<T>& y) {                                     \
    return run_stage<T, FMT<T>>(STAGE, op, alpha, mat, b, beta, y);                             \
  }
#define RVVLA_SPMV_ALL(T, FMT)                       \
  RVVLA_SPMV(T, FMT, SPARSE_MATRIX_MV_PREPARATION)   \
  RVVLA_SPMV(T, FMT, SPARSE_MATRIX_MV_MULT)          \
  RVVLA_SPMV(T, FMT, SPARSE_MATRIX_MV_GET_RESULTS)   \
  RVVLA_SPMV(T, FMT, SPARSE_MATRIX_MV_ALL_STAGES)
#define RVVLA_SPMV_FMT(FMT) RVVLA_SPMV_ALL(double, FMT) RVVLA_SPMV_ALL(float, FMT)

RVVLA_SPMV_FMT(spMtxCRS)
RVVLA_SPMV_FMT(spMtxHCSR)
RVVLA_SPMV_FMT(spMtxSELL_C_Sigma)
RVVLA_SPMV_FMT(spMtxCSR5)
RVVLA_SPMV_FMT(spMtxLAV)
RVVLA_SPMV_FMT(spMtxVNEC)
RVVLA_SPMV_FMT(spMtxVHCC)
RVVLA_SPMV_FMT(spMtxCVR)

// ---------------------------------------------------------------- convert
namespace {

int iparam(const convertParams& p, const char* k, int dflt) {
  auto it = p.param.find(k);
  return it == p.param.end() ? dflt : it->second.i;
}
double dparam(const convertParams& p, const char* k, double dflt) {
  auto it = p.param.find(k);
  return it == p.param.end() ? dflt : it->second.d;
}

}  // namespace

// merge-path partitions of at most one workgroup's LDS staging each
static int vnec_parts(int64_t items, size_t acc_bytes) {
  const int64_t cap = rvvla::vnec_items((int)acc_bytes);
  return (int)std::max<int64_t>(1, (items + cap - 1) / cap);
}

// SELL-C-sigma keys "C" (64), "sigma", "device" (1: build on the device from
// the uploaded CSR, then read the fields back; same arrays as the host build)
namespace {

template <class T>
void read_field(rvvla_mat_t A, const char* f, void* dst, size_t bytes) {
  int64_t cnt = 0;
  int eb = 0;
  const int rc = rvvla_mat_read(A, f, dst, (int64_t)bytes, &cnt, &eb);
  if (rc) throw Error(rc, rvvla_strerror(rc));
}

// The CSR's arrays copied to device memory, for the device-side builders.
template <class T>
struct DeviceCsr {
  void* rst = nullptr;
  void* col = nullptr;
  void* val = nullptr;
  explicit DeviceCsr(const spMtxCRS<T>& s) {
    rvvla::check_csr_public((int64_t)s.m, (int64_t)s.n, (int64_t)s.nz, s.Rst, s.Col);
    const size_t m = s.m, nz = s.nz;
    int rc = rvvla::hip_check(hipMalloc(&rst, (m + 1) * 4), "hipMalloc");
    if (!rc) rc = rvvla::hip_check(hipMalloc(&col, nz * 4 + 8), "hipMalloc");
    if (!rc) rc = rvvla::hip_check(hipMalloc(&val, nz * sizeof(T) + 8), "hipMalloc");
    if (!rc) rc = rvvla::hip_check(hipMemcpy(rst, s.Rst, (m + 1) * 4, hipMemcpyHostToDevice), "copy");
    if (!rc && nz) rc = rvvla::hip_check(hipMemcpy(col, s.Col, nz * 4, hipMemcpyHostToDevice), "copy");
    if (!rc && nz) rc = rvvla::hip_check(hipMemcpy(val, s.Val, nz * sizeof(T), hipMemcpyHostToDevice), "copy");
    if (rc) {
      release();
      throw Error(rc, rvvla::last_error());
    }
  }
  ~DeviceCsr() { release(); }
  void release() {
    for (void** p : {&rst, &col, &val})
      if (*p) (void)hipFree(*p), *p = nullptr;
  }
};

template <class T>
void convert_sell(spMtxSELL_C_Sigma<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int C = iparam(p, "C", 64);
  const int sigma = iparam(p, "sigma", rvvla::default_sell_sigma((int64_t)s.m, (int64_t)s.nz, s.Rst));
  if (!iparam(p, "device", 0) || C != 64) {
    rvvla::build_sell(d, s, C, sigma);
    d.device.reset(rvvla::upload_sell(dtype_of<T>(), d));
    keep_source<T>(d.device, s);
    return;
  }
  const size_t m = s.m, nz = s.nz;
  DeviceCsr<T> dc(s);
  rvvla_mat_t A = nullptr;
  int rc = rvvla_sell_from_device(dtype_of<T>(), (int64_t)m, (int64_t)s.n, (int64_t)nz,
                                  (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val, sigma,
                                  nullptr, &A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  d.device.reset(A);
  keep_source<T>(d.device, s);
  d.m = s.m;
  d.n = s.n;
  d.nz = s.nz;
  d.C = 64;
  d.Sigma = sigma;
  d.cnt_b = (m + 63) / 64;
  int64_t slots = 0;
  int eb = 0;
  rc = rvvla_mat_read(A, "vCol", nullptr, 0, &slots, &eb);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  d.vCl.resize(d.cnt_b);
  d.vCs.resize(d.cnt_b);
  d.vBs.resize(d.cnt_b);
  d.vSBs.resize(d.cnt_b);
  d.vPerm.resize(m);
  d.vCol.resize((size_t)slots);
  d.vVal.resize((size_t)slots);
  read_field<T>(A, "vCl", d.vCl.data(), d.vCl.size() * sizeof(d.vCl[0]));
  read_field<T>(A, "vCs", d.vCs.data(), d.vCs.size() * sizeof(d.vCs[0]));
  read_field<T>(A, "vBs", d.vBs.data(), d.vBs.size() * sizeof(d.vBs[0]));
  read_field<T>(A, "vSBs", d.vSBs.data(), d.vSBs.size() * sizeof(d.vSBs[0]));
  read_field<T>(A, "vPerm", d.vPerm.data(), d.vPerm.size() * sizeof(d.vPerm[0]));
  read_field<T>(A, "vCol", d.vCol.data(), d.vCol.size() * sizeof(d.vCol[0]));
  read_field<T>(A, "vVal", d.vVal.data(), d.vVal.size() * sizeof(T));
}

}  // namespace

// convert returns void as in the reference; failures are thrown as
// rvvla::Error (a std::runtime_error carrying the RVVLA_* code), where the
// reference converter throws std::invalid_argument / length_error /
// logic_error (reference src/scalar/crs_to_hcsr.cpp:38,55,87-91).
// HCSR keys "R", "C" (int) and "p" (double), defaults as the reference's
// (reference src/scalar/crs_to_hcsr.cpp:141-149, :158-166), and "device"
// (1: build on the device from the uploaded CSR, then read the reference
// fields back; the same arrays as the host build).
namespace {

template <class T>
void convert_hcsr(spMtxHCSR<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int dflt = std::is_same<T, double>::value ? 2000 : 4000;
  const int64_t R = iparam(p, "R", dflt), C = iparam(p, "C", dflt);
  const double pp = dparam(p, "p", 2.0);
  if (!iparam(p, "device", 0)) {
    rvvla::build_hcsr<T>(d, s, (size_t)R, (size_t)C, pp);
    d.device.reset(rvvla::upload_hcsr<T>(dtype_of<T>(), d));
    keep_source<T>(d.device, s);
    return;
  }
  if (R <= 0 || C <= 0) throw Error(RVVLA_ERR_INVALID_ARG, "HCSR: R and C must be positive");
  DeviceCsr<T> dc(s);
  rvvla_mat_t A = nullptr;
  int rc = rvvla_hcsr_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                  (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val, R, C, pp,
                                  nullptr, &A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  d = spMtxHCSR<T>();
  d.device.reset(A);
  keep_source<T>(d.device, s);
  d.R = (size_t)R;
  d.C = (size_t)C;
  d.m = s.m;
  d.n = s.n;
  d.nz = s.nz;
  d.b_m = (s.m + (size_t)R - 1) / (size_t)R;
  d.b_n = (s.n + (size_t)C - 1) / (size_t)C;
  int64_t nrp = 0;
  int eb = 0;
  rc = rvvla_mat_read(A, "row_ptr", nullptr, 0, &nrp, &eb);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  const size_t nb = d.b_m * d.b_n + 1;
  d.block_ptr.resize(nb);
  d.row_offset.resize(nb);
  d.row_ptr.resize((size_t)nrp);
  d.col_idx.resize(s.nz);
  d.values.resize(s.nz);
  read_field<T>(A, "block_ptr", d.block_ptr.data(), nb * 4);
  read_field<T>(A, "row_offset", d.row_offset.data(), nb * 4);
  read_field<T>(A, "row_ptr", d.row_ptr.data(), (size_t)nrp * 4);
  read_field<T>(A, "col_idx", d.col_idx.data(), s.nz * 2);
  read_field<T>(A, "values", d.values.data(), s.nz * sizeof(T));
}

}  // namespace

template <>
void convert<double, spMtxHCSR, true>(spMtxHCSR<double>& d, const spMtxCRS<double>& s,
                                      const convertParams& p) {
  convert_hcsr<double>(d, s, p);
}
template <>
void convert<float, spMtxHCSR, true>(spMtxHCSR<float>& d, const spMtxCRS<float>& s,
                                     const convertParams& p) {
  convert_hcsr<float>(d, s, p);
}

template <>
void convert<double, spMtxSELL_C_Sigma, true>(spMtxSELL_C_Sigma<double>& d,
                                              const spMtxCRS<double>& s, const convertParams& p) {
  convert_sell<double>(d, s, p);
}
template <>
void convert<float, spMtxSELL_C_Sigma, true>(spMtxSELL_C_Sigma<float>& d,
                                             const spMtxCRS<float>& s, const convertParams& p) {
  convert_sell<float>(d, s, p);
}
// The remaining formats take "device" = 1 as SELL-C-sigma and HCSR do:
// build on the device from the uploaded CSR, keep that image for the
// products, and fill the container from it (the same arrays as the host
// build).
namespace {

template <class T>
int64_t field_count(rvvla_mat_t A, const char* f) {
  int64_t cnt = 0;
  int eb = 0;
  const int rc = rvvla_mat_read(A, f, nullptr, 0, &cnt, &eb);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  return cnt;
}

template <class V, class T>
std::vector<V> read_vec(rvvla_mat_t A, const char* f) {
  std::vector<V> v((size_t)field_count<T>(A, f));
  if (!v.empty()) read_field<T>(A, f, v.data(), v.size() * sizeof(V));
  return v;
}

struct MatHandle {
  rvvla_mat_t A = nullptr;
  ~MatHandle() {
    if (A) (void)rvvla_mat_free(A);
  }
  rvvla_mat_t release() {
    rvvla_mat_t a = A;
    A = nullptr;
    return a;
  }
};

// CSR5 keys "omega" (64), "sigma" (0 = chosen from the row lengths)
template <class T>
void convert_csr5(spMtxCSR5<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int omega = iparam(p, "omega", 64);
  int sigma = iparam(p, "sigma", 0);
  if (!iparam(p, "device", 0)) {
    rvvla::build_csr5(d, s, omega, sigma);
    d.device.reset(rvvla::upload_csr5(dtype_of<T>(), d));
    keep_source<T>(d.device, s);
    return;
  }
  if (omega != 64) throw Error(RVVLA_ERR_UNSUPPORTED_OMEGA, "CSR5: omega must be 64");
  if (sigma <= 0) sigma = rvvla::default_csr5_sigma((int64_t)s.m, (int64_t)s.nz);
  if (sigma > 64) throw Error(RVVLA_ERR_CONVERSION, "CSR5: sigma must be in [1, 64]");
  DeviceCsr<T> dc(s);
  MatHandle H;
  int rc = rvvla_csr5_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                  (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val, sigma,
                                  nullptr, &H.A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  d.freeMem();
  rc = d.inputCSR((int)s.m, (int)s.n, (int)s.nz, s.Rst, s.Col, s.Val);
  if (rc) throw Error(RVVLA_ERR_CONVERSION, "CSR5: inputCSR failed");
  const int64_t nnz = (int64_t)s.nz, W = 64 * (int64_t)sigma;
  d._omega = 64;
  d._sigma = sigma;
  d._p = nnz > 0 ? (int)((nnz + W - 1) / W) : 1;
  int by = 0, bs = 0;
  while ((int64_t(1) << by) < W) ++by;
  while ((1 << bs) < 64) ++bs;
  d._bit_y_offset = by;
  d._bit_seg_offset = bs;
  d._tile_desc_len = (by + bs + sigma + 31) / 32;
  d._csr5_tile_pointer = read_vec<uint32_t, T>(H.A, "_csr5_tile_pointer");
  d._csr5_tile_descriptor = read_vec<uint32_t, T>(H.A, "_csr5_tile_descriptor");
  d._csr5_tile_empty_offset_pointer = read_vec<uint32_t, T>(H.A, "_csr5_tile_empty_offset_pointer");
  d._csr5_tile_empty_offset = read_vec<uint32_t, T>(H.A, "_csr5_tile_empty_offset");
  d._num_offsets = (int)d._csr5_tile_empty_offset.size();
  d._csr_column_index = read_vec<uint32_t, T>(H.A, "_csr_column_index");
  d._csr_value = read_vec<T, T>(H.A, "_csr_value");
  d._tail_tile_start = (int)(d._csr5_tile_pointer[(size_t)d._p - 1] & 0x7FFFFFFFu);
  d._temp_calibrator.assign((size_t)omp_get_max_threads() * (CSR5_CACHELINE / sizeof(T)), T(0));
  d.setCSR5Format();
  d.device.reset(H.release());
  keep_source<T>(d.device, s);
}

// LAV keys "SIMD_Lanes" (64), "T" (double, 0.2), "segment_len"
template <class T>
void convert_lav(spMtxLAV<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int lanes = iparam(p, "SIMD_Lanes", 64);
  const double Tv = dparam(p, "T", 0.2);
  const int seg = iparam(p, "segment_len", 1 << 20);
  if (!iparam(p, "device", 0) || lanes != 64) {
    rvvla::build_lav(d, s, lanes, Tv, seg);
    d.device.reset(rvvla::upload_lav(dtype_of<T>(), d));
    keep_source<T>(d.device, s);
    return;
  }
  if (seg < 1) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: segment_len must be positive");
  if (!(Tv >= 0.0 && Tv <= 1.0)) throw Error(RVVLA_ERR_INVALID_ARG, "LAV: T in [0, 1]");
  DeviceCsr<T> dc(s);
  MatHandle H;
  const int rc = rvvla_lav_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                       (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val, Tv,
                                       seg, nullptr, &H.A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  d = spMtxLAV<T>();
  d.SIMD_Lanes = 64;
  d.m = (int)s.m;
  d.n = (int)s.n;
  d.nz = (int)s.nz;
  d.T = Tv;
  d.segment_len = seg;
  d.chunk_count = (s.m + 63) / 64;
  d.chunk_offsets = read_vec<int, T>(H.A, "chunk_offsets");
  d.chunk_offsets_full_mask = read_vec<int, T>(H.A, "chunk_offsets_full_mask");
  d.segment_ptr = read_vec<size_t, T>(H.A, "segment_ptr");
  d.segment_count = d.segment_ptr.size() - 1;
  d.out_order = read_vec<uint32_t, T>(H.A, "out_order");
  d.mask = read_vec<uint16_t, T>(H.A, "mask");
  d.val = read_vec<T, T>(H.A, "val");
  d.col_id = read_vec<uint32_t, T>(H.A, "col_id");
  const std::vector<int> srst = read_vec<int, T>(H.A, "sparse_part.Rst");
  const std::vector<int> scol = read_vec<int, T>(H.A, "sparse_part.Col");
  const std::vector<T> sval = read_vec<T, T>(H.A, "sparse_part.Val");
  d.sparse_part = spMtxCRS<T>(s.m, s.n, scol.size(), srst.data(), scol.data(), sval.data());
  d.device.reset(H.release());
  keep_source<T>(d.device, s);
}

// CVR keys "SIMD_LEN" (64), "streams"
template <class T>
void convert_cvr(spMtxCVR<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int L = iparam(p, "SIMD_LEN", 64);
  const int S = iparam(p, "streams",
                       rvvla::default_parts((int64_t)s.m, (int64_t)s.nz, rvvla::kCvrStreamEntries));
  if (!iparam(p, "device", 0) || L != 64) {
    rvvla::build_cvr(d, s, L, S);
    d.device.reset(rvvla::upload_cvr(dtype_of<T>(), d, (int64_t)s.nz));
    keep_source<T>(d.device, s);
    return;
  }
  if (S < 1) throw Error(RVVLA_ERR_INVALID_ARG, "CVR: at least one stream");
  DeviceCsr<T> dc(s);
  MatHandle H;
  const int rc = rvvla_cvr_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                       (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val, S,
                                       nullptr, &H.A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  const std::vector<int64_t> ps = read_vec<int64_t, T>(H.A, "pack_start");
  const std::vector<uint64_t> flush = read_vec<uint64_t, T>(H.A, "flush");
  const std::vector<int32_t> rec_rows = read_vec<int32_t, T>(H.A, "rec_rows");
  const std::vector<int32_t> tails = read_vec<int32_t, T>(H.A, "tail_rows");
  // flush targets of cut rows name a partial-sum slot: its row
  const std::vector<int32_t> srows = read_vec<int32_t, T>(H.A, "split_rows");
  const std::vector<int32_t> sptr = read_vec<int32_t, T>(H.A, "split_ptr");
  std::vector<int32_t> slot_row;
  for (size_t i = 0; i < srows.size(); ++i) {
    if ((size_t)sptr[2 * i + 1] > slot_row.size()) slot_row.resize((size_t)sptr[2 * i + 1], -1);
    for (int32_t q = sptr[2 * i]; q < sptr[2 * i + 1]; ++q) slot_row[(size_t)q] = srows[i];
  }
  auto row_of = [&](int32_t target) { return target >= 0 ? target : slot_row[(size_t)(-target - 2)]; };
  d = spMtxCVR<T>();
  d.SIMD_LEN = 64;
  d.numRows = s.m;
  d.numCols = s.n;
  d.Nthreads = d.Nthreads_avail = S;
  d.vPack_vec_vals = read_vec<T, T>(H.A, "val");
  d.vPack_vec_cols = read_vec<uint32_t, T>(H.A, "col");
  d.nItems = d.vPack_vec_vals.size();
  // records: per stream the lanes of each pack's flush mask in lane order,
  // then the rows flushed after its last pack
  d.writeback_pos_record.assign((size_t)S * 66, -1);
  size_t q = 0;
  for (int t = 0; t < S; ++t) {
    int* wb = d.writeback_pos_record.data() + (size_t)t * 66;
    wb[0] = (int)ps[(size_t)t];
    wb[1] = (int)(d.vPack_vec_record.size() / 2);
    for (int64_t k = ps[(size_t)t]; k < ps[(size_t)t + 1]; ++k)
      for (uint64_t f = flush[(size_t)k]; f; f &= f - 1) {
        d.vPack_vec_record.push_back((int)(k * 64 + __builtin_ctzll(f)));
        d.vPack_vec_record.push_back(row_of(rec_rows[q++]));
      }
    for (int l = 0; l < 64; ++l)
      if (tails[(size_t)t * 64 + l] != -1) {
        d.vPack_vec_record.push_back((int)(ps[(size_t)t + 1] * 64 + l));
        d.vPack_vec_record.push_back(row_of(tails[(size_t)t * 64 + l]));
      }
  }
  d.vPack_vec_record.push_back((int)d.nItems);
  d.vPack_vec_record.push_back(-1);
  if (S == 1) {  // single stream: idle lanes name the last non-empty row
    int last = -1;
    for (int64_t r = (int64_t)s.m - 1; r >= 0; --r)
      if (s.Rst[r + 1] > s.Rst[r]) {
        last = (int)r;
        break;
      }
    for (int l = 0; l < 64; ++l) d.writeback_pos_record[2 + (size_t)l] = last < 0 ? 0 : last;
  }
  d.device.reset(H.release());
  keep_source<T>(d.device, s);
}

// VHCC keys "vec_size" (64), "threads"
template <class T>
void convert_vhcc(spMtxVHCC<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int vec = iparam(p, "vec_size", 64);
  const int threads = iparam(p, "threads", rvvla::default_parts((int64_t)s.m, (int64_t)s.nz,
                                                                rvvla::kVhccPartEntries));
  if (!iparam(p, "device", 0) || vec != 64) {
    rvvla::build_vhcc(d, s, vec, threads);
    d.device.reset(rvvla::upload_vhcc(dtype_of<T>(), d));
    keep_source<T>(d.device, s);
    return;
  }
  if (threads < 1) throw Error(RVVLA_ERR_INVALID_ARG, "VHCC: at least one partition");
  DeviceCsr<T> dc(s);
  MatHandle H;
  const int rc = rvvla_vhcc_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                        (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val,
                                        threads, nullptr, &H.A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  const std::vector<int32_t> sv = read_vec<int32_t, T>(H.A, "start_vec");
  const std::vector<int32_t> rbase = read_vec<int32_t, T>(H.A, "rbase");
  const std::vector<uint64_t> eor = read_vec<uint64_t, T>(H.A, "eor");
  const std::vector<int32_t> fwd = read_vec<int32_t, T>(H.A, "fwd_row");
  const std::vector<uint8_t> pass = read_vec<uint8_t, T>(H.A, "pass");
  const int64_t m = (int64_t)s.m, nnz = (int64_t)s.nz, Tn = (int64_t)rbase.size();
  const int64_t nvec = (int64_t)eor.size();
  d.freeMem();
  d.set_params(vhcc_params((size_t)Tn, 1, 512, 512, 64));
  d.set_properties((size_t)m, s.n, (size_t)nnz, (size_t)nvec, (size_t)(nvec * 64 - nnz),
                   (size_t)(m + Tn));
  d.get_vals_arr_reference() = read_vec<T, T>(H.A, "val");
  d.get_col_arr_reference() = read_vec<uint32_t, T>(H.A, "col");
  d.get_row_arr_reference() = read_vec<uint32_t, T>(H.A, "row_arr");
  // row-end vectors and their masks: lane j adds lane j - 2^st in step st
  // when no row ends in lanes [j - 2^st, j); the write mask is the row ends
  auto& veceor = d.get_veceor_reference();
  auto& masks = d.get_scan_mask_reference();
  auto& thr = d.get_thr_info_reference();
  thr.assign((size_t)Tn, thr_info_t{});
  int last_row = -1;
  for (int64_t t = 0; t < Tn; ++t) {
    thr_info_t& I = thr[(size_t)t];
    I.start_vec = sv[(size_t)t];
    I.end_vec = sv[(size_t)t + 1];
    I.panel_id = 0;
    I.merge_start = (int)std::min<int64_t>((int64_t)I.start_vec * 64, nnz);
    I.merge_end = (int)std::min<int64_t>((int64_t)I.end_vec * 64, nnz);
    I.last_row = I.overflow_row = (int)(m + t);
    // partition t continues a row iff t - 1 hands its partial on
    if (t > 0 && (fwd[(size_t)t - 1] >= 0 || pass[(size_t)t - 1])) {
      last_row = fwd[(size_t)t - 1] >= 0 ? fwd[(size_t)t - 1] : last_row;
      I.last_row = last_row;
    }
    I.vbase = (int)veceor.size();
    I.rbase = rbase[(size_t)t];
    I.sbase = (int)(veceor.size() * 7);
    for (int64_t v = I.start_vec; v < I.end_vec; ++v) {
      const uint64_t E = eor[(size_t)v];
      if (!E) continue;
      veceor.push_back((uint32_t)v);
      uint64_t G = E << 1;  // bit j: a row ends in lanes [j - span, j)
      for (int st = 0; st <= 6; ++st) {
        uint64_t mk = E;
        if (st < 6) {
          if (st > 0) G |= G << (1 << (st - 1));
          mk = ~G & ~((uint64_t(1) << (1 << st)) - 1);
        }
        for (int b = 0; b < 8; ++b) masks.push_back((uint8_t)(mk >> (8 * b)));
      }
    }
  }
  veceor.push_back((uint32_t)nvec);
  d.device.reset(H.release());
  keep_source<T>(d.device, s);
}

// VNEC key "threads" (merge-path partitions; default: as many rows+entries
// each as one workgroup stages, 1024 for double, 2048 for float)
template <class T>
void convert_vnec(spMtxVNEC<T>& d, const spMtxCRS<T>& s, const convertParams& p) {
  const int nparts =
      iparam(p, "threads", vnec_parts((int64_t)s.m + (int64_t)s.nz, sizeof(*s.Val)));
  if (!iparam(p, "device", 0)) {
    rvvla::build_vnec(d, s, nparts);
    d.device.reset(rvvla::upload_vnec(dtype_of<T>(), d, s));
    keep_source<T>(d.device, s);
    return;
  }
  if (nparts < 1) throw Error(RVVLA_ERR_INVALID_ARG, "VNEC: at least one partition");
  DeviceCsr<T> dc(s);
  MatHandle H;
  const int rc = rvvla_vnec_from_device(dtype_of<T>(), (int64_t)s.m, (int64_t)s.n, (int64_t)s.nz,
                                        (const int32_t*)dc.rst, (const int32_t*)dc.col, dc.val,
                                        nparts, nullptr, &H.A);
  if (rc) throw Error(rc, rvvla_strerror(rc));
  const std::vector<int32_t> cx = read_vec<int32_t, T>(H.A, "cx");
  const std::vector<int32_t> cy = read_vec<int32_t, T>(H.A, "cy");
  const std::vector<uint16_t> ecr = read_vec<uint16_t, T>(H.A, "ecr_indices");
  d = spMtxVNEC<T>();
  d.G = &s;
  d.m = (int)s.m;
  d.n = (int)s.n;
  d.nnz = (int)s.nz;
  d.NUM_THREADS_VNEC = nparts;
  d.thread_coord_start.resize((size_t)nparts);
  d.thread_coord_end.resize((size_t)nparts);
  for (int t = 0; t < nparts; ++t) {
    d.thread_coord_start[(size_t)t] = coord{cx[(size_t)t], cy[(size_t)t]};
    d.thread_coord_end[(size_t)t] = coord{cx[(size_t)t + 1], cy[(size_t)t + 1]};
  }
  d.ecr_indices.assign(ecr.begin(), ecr.end());
  d.nec_ptr = read_vec<int, T>(H.A, "nec_ptr");
  d.NEC_NUM.resize((size_t)nparts);
  for (int t = 0; t < nparts; ++t) d.NEC_NUM[(size_t)t] = d.nec_ptr[(size_t)t + 1] - d.nec_ptr[(size_t)t];
  d.use_x_indices = read_vec<int, T>(H.A, "use_x_indices");
  rvvla::vnec_select_variant(d, s);
  d.device.reset(H.release());
  keep_source<T>(d.device, s);
}

}  // namespace

#define RVVLA_CONVERT_FN(FMT, FN)                                                                 \
  template <>                                                                                     \
  void convert<double, FMT, true>(FMT<double> & d, const spMtxCRS<double>& s,                     \
                                  const convertParams& p) {                                       \
    FN<double>(d, s, p);                                                                          \
  }                                                                                               \
  template <>                                                                                     \
  void convert<float, FMT, true>(FMT<float> & d, const spMtxCRS<float>& s,                        \
                                 const convertParams& p) {                                        \
    FN<float>(d, s, p);                                                                           \
  }

RVVLA_CONVERT_FN(spMtxCSR5, convert_csr5)
RVVLA_CONVERT_FN(spMtxLAV, convert_lav)
RVVLA_CONVERT_FN(spMtxCVR, convert_cvr)
RVVLA_CONVERT_FN(spMtxVHCC, convert_vhcc)
RVVLA_CONVERT_FN(spMtxVNEC, convert_vnec)

}  // namespace SparseMatrixLib
