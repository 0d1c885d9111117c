// cxx_api.cpp -- the C++ drop-in: convert<T, Fmt, true> and
// sparse_mv<T, Fmt, true, stage> for the eight formats, on top of the C-ABI.
//
// Mirrors the reference API (reference include/spmv_mtx.h:21-29,
// include/convert.h:30-31): the same template names, arguments and
// std::vector signature.  convert builds the format's host fields and
// uploads its device image once (dist.device); sparse_mv uploads lazily when
// a container was filled some other way.  Stages (reference
// include/types.h:33-38): PREPARATION copies b and y to the device,
// MULT runs the kernel, GET_RESULTS copies y back, ALL_STAGES does all
// three.  Status codes are the RVVLA_* codes (0 = success).  The reference
// ignores type_op; here SPARSE_OPERATION_TRANSPOSE computes A^T b for CRS
// matrices and is answered with RVVLA_ERR_UNSUPPORTED_SPMV by the other
// formats instead of computing the untransposed product.
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

template <class M>
constexpr bool is_crs() {
  return false;
}
template <>
constexpr bool is_crs<spMtxCRS<double>>() {
  return true;
}
template <>
constexpr bool is_crs<spMtxCRS<float>>() {
  return true;
}

template <class T, class M>
sparse_matrix_status run_stage(sparse_matrix_mv_stage st, sparse_operation_t op, T alpha, const M& mat,
                               const std::vector<T>& b, T beta, std::vector<T>& y) {
  sparse_matrix_status status;
  if (op != SPARSE_OPERATION_NON_TRANSPOSE && !(op == SPARSE_OPERATION_TRANSPOSE && is_crs<M>())) {
    rvvla::set_last_error("SPARSE_OPERATION_TRANSPOSE is implemented for CRS matrices");
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

#define RVVLA_CONVERT(T, FMT, BUILD, UPLOAD)                                                       \
  template <>                                                                                      \
  void convert<T, FMT, true>(FMT<T> & d, const spMtxCRS<T>& s, const convertParams& p) {           \
    BUILD;                                                                                         \
    d.device.reset(UPLOAD);                                                                        \
  }

// merge-path partitions of at most one workgroup's LDS staging each
static int vnec_parts(int64_t items, size_t acc_bytes) {
  const int64_t cap = rvvla::vnec_items((int)acc_bytes);
  return (int)std::max<int64_t>(1, (items + cap - 1) / cap);
}

#define RVVLA_CONVERT_BOTH(FMT, BUILD, UPLOAD)                                                     \
  RVVLA_CONVERT(double, FMT, BUILD, UPLOAD) RVVLA_CONVERT(float, FMT, BUILD, UPLOAD)

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
// CSR5 keys "omega" (64), "sigma" (0 = chosen from the row lengths)
RVVLA_CONVERT_BOTH(spMtxCSR5, rvvla::build_csr5(d, s, iparam(p, "omega", 64), iparam(p, "sigma", 0)),
                   rvvla::upload_csr5(dtype_of<std::decay_t<decltype(*s.Val)>>(), d))
// LAV keys "SIMD_Lanes" (64), "T" (double, 0.2), "segment_len"
RVVLA_CONVERT_BOTH(spMtxLAV,
                   rvvla::build_lav(d, s, iparam(p, "SIMD_Lanes", 64), dparam(p, "T", 0.2),
                                    iparam(p, "segment_len", 1 << 20)),
                   rvvla::upload_lav(dtype_of<std::decay_t<decltype(*s.Val)>>(), d))
// CVR keys "SIMD_LEN" (64), "streams"
RVVLA_CONVERT_BOTH(spMtxCVR,
                   rvvla::build_cvr(d, s, iparam(p, "SIMD_LEN", 64),
                                    iparam(p, "streams", rvvla::default_parts((int64_t)s.m, (int64_t)s.nz, rvvla::kCvrStreamEntries))),
                   rvvla::upload_cvr(dtype_of<std::decay_t<decltype(*s.Val)>>(), d, (int64_t)s.nz))
// VHCC keys "vec_size" (64), "threads"
RVVLA_CONVERT_BOTH(spMtxVHCC,
                   rvvla::build_vhcc(d, s, iparam(p, "vec_size", 64),
                                     iparam(p, "threads", rvvla::default_parts((int64_t)s.m, (int64_t)s.nz, rvvla::kVhccPartEntries))),
                   rvvla::upload_vhcc(dtype_of<std::decay_t<decltype(*s.Val)>>(), d))
// VNEC key "threads" (merge-path partitions; default: as many rows+entries
// each as one workgroup stages, 1024 for double, 2048 for float)
RVVLA_CONVERT_BOTH(spMtxVNEC,
                   rvvla::build_vnec(d, s, iparam(p, "threads", vnec_parts((int64_t)s.m + (int64_t)s.nz, sizeof(*s.Val)))),
                   rvvla::upload_vnec(dtype_of<std::decay_t<decltype(*s.Val)>>(), d, s))

}  // namespace SparseMatrixLib
