// ref_shim.cpp -- C entry points over the reference's own scalar library,
// compiled from the reference sources where they lie (see Makefile).  TEST
// INFRASTRUCTURE ONLY: it lets the tests run the real reference converter
// and scalar kernels on the same inputs as the oracle restatement.  Nothing
// here is shipped or loaded by the library.
#include <cstdint>
#include <cstring>
#include <vector>

#include "sparse_matrix.h"

using namespace SparseMatrixLib;

namespace {

template <class T>
spMtxCRS<T> make_crs(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                     const T* val) {
  return spMtxCRS<T>((size_t)m, (size_t)n, (size_t)nnz, rst, col, val);
}

template <class T>
struct HcsrBox {
  spMtxHCSR<T> h;
};

struct RefBenchBase {
  virtual ~RefBenchBase() = default;
  virtual void run() = 0;
  virtual const void* y() const = 0;
};
template <class T, template <class> class Fmt>
struct RefBench : RefBenchBase {
  Fmt<T> a;
  std::vector<T> b, yv;
  void run() override {
    sparse_mv<T, Fmt, false>(SPARSE_OPERATION_NON_TRANSPOSE, (T)1, a, sparse_matrix_descr{}, b,
                             (T)0, yv);
  }
  const void* y() const override { return yv.data(); }
};

}  // namespace

extern "C" {

// reference src/scalar/spmv_mtx_crs.cpp:22-71
int ref_csr_spmv_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                     const double* val, const double* b, double alpha, double beta, double* y) {
  auto a = make_crs<double>(m, n, nnz, rst, col, val);
  std::vector<double> vb(b, b + n), vy(y, y + m);
  sparse_mv<double, spMtxCRS, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, a,
                                     sparse_matrix_descr{}, vb, beta, vy);
  std::memcpy(y, vy.data(), sizeof(double) * m);
  return 0;
}
int ref_csr_spmv_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                     const float* val, const float* b, float alpha, float beta, float* y) {
  auto a = make_crs<float>(m, n, nnz, rst, col, val);
  std::vector<float> vb(b, b + n), vy(y, y + m);
  sparse_mv<float, spMtxCRS, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, a,
                                    sparse_matrix_descr{}, vb, beta, vy);
  std::memcpy(y, vy.data(), sizeof(float) * m);
  return 0;
}

// The reference CSR and HCSR SpMVs on a matrix and vectors built once, for
// timing (bench.py's CPU baselines): new (copies the matrix and b in; HCSR:
// converts through convert<T, spMtxHCSR, false> with R, C, p), run (one
// y = A*b through sparse_mv<T, Fmt, false>, reference
// src/scalar/spmv_mtx_crs.cpp:22-71 and src/scalar/spmv_mtx_hcsr.cpp
// :104-134), y (the last result), free.
void* ref_csr_bench_new_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                            const double* val, const double* b) {
  auto* p = new RefBench<double, spMtxCRS>;
  p->a = make_crs<double>(m, n, nnz, rst, col, val);
  p->b.assign(b, b + n);
  p->yv.assign((size_t)m, 0.0);
  return static_cast<RefBenchBase*>(p);
}
void* ref_csr_bench_new_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                            const float* val, const float* b) {
  auto* p = new RefBench<float, spMtxCRS>;
  p->a = make_crs<float>(m, n, nnz, rst, col, val);
  p->b.assign(b, b + n);
  p->yv.assign((size_t)m, 0.0f);
  return static_cast<RefBenchBase*>(p);
}
void* ref_hcsr_bench_new_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                             const double* val, const double* b, int R, int C, double p) {
  auto* q = new RefBench<double, spMtxHCSR>;
  convertParams P;
  P.param["R"].i = R;
  P.param["C"].i = C;
  P.param["p"].d = p;
  auto crs = make_crs<double>(m, n, nnz, rst, col, val);
  convert<double, spMtxHCSR, false>(q->a, crs, P);
  q->b.assign(b, b + n);
  q->yv.assign((size_t)m, 0.0);
  return static_cast<RefBenchBase*>(q);
}
int ref_csr_bench_run_f64(void* h) {
  static_cast<RefBenchBase*>(h)->run();
  return 0;
}
int ref_bench_run(void* h) {
  static_cast<RefBenchBase*>(h)->run();
  return 0;
}
const void* ref_bench_y(void* h) { return static_cast<RefBenchBase*>(h)->y(); }
void ref_csr_bench_free(void* h) { delete static_cast<RefBenchBase*>(h); }

// reference src/scalar/crs_to_hcsr.cpp:137-169 through convert<T, spMtxHCSR, false>.
// Returns a handle; *err = 0, or -10/-11/-12 for the three exception kinds.
static void* hcsr_convert_any(int dtype, int64_t m, int64_t n, int64_t nnz, const int* rst,
                              const int* col, const void* val, int R, int C, double p, int* err) {
  convertParams P;
  P.param["R"].i = R;
  P.param["C"].i = C;
  P.param["p"].d = p;
  *err = 0;
  try {
    if (dtype == 0) {
      auto* box = new HcsrBox<double>;
      auto a = make_crs<double>(m, n, nnz, rst, col, (const double*)val);
      convert<double, spMtxHCSR, false>(box->h, a, P);
      return box;
    }
    auto* box = new HcsrBox<float>;
    auto a = make_crs<float>(m, n, nnz, rst, col, (const float*)val);
    convert<float, spMtxHCSR, false>(box->h, a, P);
    return box;
  } catch (const std::invalid_argument&) {
    *err = -10;
  } catch (const std::length_error&) {
    *err = -11;
  } catch (const std::logic_error&) {
    *err = -12;
  }
  return nullptr;
}
void* ref_hcsr_convert_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                           const double* val, int R, int C, double p, int* err) {
  return hcsr_convert_any(0, m, n, nnz, rst, col, val, R, C, p, err);
}
void* ref_hcsr_convert_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                           const float* val, int R, int C, double p, int* err) {
  return hcsr_convert_any(1, m, n, nnz, rst, col, val, R, C, p, err);
}

}  // extern "C"

template <class T>
static int hcsr_get(void* hb, const char* f, const void** data, int64_t* count, int* elem) {
  auto& h = ((HcsrBox<T>*)hb)->h;
  auto put = [&](const auto& v) {
    *data = v.data();
    *count = (int64_t)v.size();
    *elem = (int)sizeof(v[0]);
    return 0;
  };
  if (!std::strcmp(f, "block_ptr")) return put(h.block_ptr);
  if (!std::strcmp(f, "row_offset")) return put(h.row_offset);
  if (!std::strcmp(f, "row_ptr")) return put(h.row_ptr);
  if (!std::strcmp(f, "col_idx")) return put(h.col_idx);
  if (!std::strcmp(f, "values")) return put(h.values);
  return -23;
}

extern "C" {
int ref_hcsr_array_f64(void* h, const char* f, const void** d, int64_t* c, int* e) {
  return hcsr_get<double>(h, f, d, c, e);
}
int ref_hcsr_array_f32(void* h, const char* f, const void** d, int64_t* c, int* e) {
  return hcsr_get<float>(h, f, d, c, e);
}
int ref_hcsr_dims_f64(void* hb, int64_t* out7) {
  auto& h = ((HcsrBox<double>*)hb)->h;
  int64_t v[7] = {(int64_t)h.R, (int64_t)h.C, (int64_t)h.m, (int64_t)h.n, (int64_t)h.nz,
                  (int64_t)h.b_m, (int64_t)h.b_n};
  std::memcpy(out7, v, sizeof v);
  return 0;
}
int ref_hcsr_dims_f32(void* hb, int64_t* out7) {
  auto& h = ((HcsrBox<float>*)hb)->h;
  int64_t v[7] = {(int64_t)h.R, (int64_t)h.C, (int64_t)h.m, (int64_t)h.n, (int64_t)h.nz,
                  (int64_t)h.b_m, (int64_t)h.b_n};
  std::memcpy(out7, v, sizeof v);
  return 0;
}
// reference src/scalar/spmv_mtx_hcsr.cpp:104-134
int ref_hcsr_spmv_f64(void* hb, const double* b, double alpha, double beta, double* y) {
  auto& h = ((HcsrBox<double>*)hb)->h;
  std::vector<double> vb(b, b + h.n), vy(y, y + h.m);
  sparse_mv<double, spMtxHCSR, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, h,
                                      sparse_matrix_descr{}, vb, beta, vy);
  std::memcpy(y, vy.data(), sizeof(double) * h.m);
  return 0;
}
int ref_hcsr_spmv_f32(void* hb, const float* b, float alpha, float beta, float* y) {
  auto& h = ((HcsrBox<float>*)hb)->h;
  std::vector<float> vb(b, b + h.n), vy(y, y + h.m);
  sparse_mv<float, spMtxHCSR, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, h,
                                     sparse_matrix_descr{}, vb, beta, vy);
  std::memcpy(y, vy.data(), sizeof(float) * h.m);
  return 0;
}
void ref_hcsr_free_f64(void* h) { delete (HcsrBox<double>*)h; }
void ref_hcsr_free_f32(void* h) { delete (HcsrBox<float>*)h; }

// reference src/scalar/spmv_mtx_sell_c_sigma.cpp:130-149 (ALL_STAGES) on
// caller-provided SELL arrays.
}  // extern "C"

template <class T>
static int sell_run(int64_t m, int64_t n, int64_t nb, int C, int sigma, const int* Cl,
                    const int* Cs, const int* Bs, const int* SBs, int64_t slots,
                    const uint32_t* Col, const uint32_t* Perm, const T* Val, const T* b, T alpha,
                    T beta, T* y) {
  spMtxSELL_C_Sigma<T> s;
  s.m = (size_t)m;
  s.n = (size_t)n;
  s.cnt_b = (size_t)nb;
  s.C = C;
  s.Sigma = sigma;
  s.vCl.assign(Cl, Cl + nb);
  s.vCs.assign(Cs, Cs + nb);
  s.vBs.assign(Bs, Bs + nb);
  s.vSBs.assign(SBs, SBs + nb);
  s.vCol.assign(Col, Col + slots);
  s.vPerm.assign(Perm, Perm + m);
  s.vVal.assign(Val, Val + slots);
  std::vector<T> vb(b, b + n), vy(y, y + m);
  sparse_mv<T, spMtxSELL_C_Sigma, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, s,
                                         sparse_matrix_descr{}, vb, beta, vy);
  std::memcpy(y, vy.data(), sizeof(T) * m);
  return 0;
}

// reference src/scalar/spmv_mtx_cvr.cpp:21-118 on caller-provided CVR arrays.
template <class T>
static int cvr_run(int64_t m, int64_t n, int64_t nitems, int simd_len, int nthreads,
                   const T* vals, const uint32_t* cols, int64_t nrec2, const int* rec,
                   const int* wb, const T* b, T alpha, T beta, T* y) {
  spMtxCVR<T> c;
  c.nItems = (size_t)nitems;
  c.numRows = (size_t)m;
  c.numCols = (size_t)n;
  c.SIMD_LEN = (size_t)simd_len;
  c.Nthreads = c.Nthreads_avail = nthreads;
  c.vPack_vec_vals.assign(vals, vals + nitems);
  c.vPack_vec_cols.assign(cols, cols + nitems);
  c.vPack_vec_record.assign(rec, rec + nrec2);
  c.writeback_pos_record.assign(wb, wb + (size_t)nthreads * (simd_len + 2));
  std::vector<T> vb(b, b + n), vy(y, y + m);
  sparse_mv<T, spMtxCVR, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, c, sparse_matrix_descr{}, vb,
                                beta, vy);
  std::memcpy(y, vy.data(), sizeof(T) * m);
  return 0;
}

// reference src/scalar/spmv_mtx_lav.cpp:23-144 on caller-provided LAV arrays
// (the sparse part as CSR arrays).
template <class T>
static int lav_run(int64_t m, int64_t n, int64_t nz, int lanes, int64_t segment_count,
                   int64_t chunk_count, const int* chunk_offsets, int64_t n_co,
                   const int* chunk_full, int64_t n_cf, const uint64_t* segment_ptr, int64_t n_sp,
                   const uint32_t* out_order, int64_t n_oo, const uint16_t* mask, int64_t n_mask,
                   const T* val, const uint32_t* col_id, int64_t n_val, int64_t sp_nz,
                   const int* sp_rst, const int* sp_col, const T* sp_val, const T* b, T alpha,
                   T beta, T* y) {
  spMtxLAV<T> l;
  l.SIMD_Lanes = (size_t)lanes;
  l.m = (int)m;
  l.n = (int)n;
  l.nz = (int)nz;
  l.segment_count = (size_t)segment_count;
  l.chunk_count = (size_t)chunk_count;
  l.chunk_offsets.assign(chunk_offsets, chunk_offsets + n_co);
  l.chunk_offsets_full_mask.assign(chunk_full, chunk_full + n_cf);
  l.segment_ptr.assign(segment_ptr, segment_ptr + n_sp);
  l.out_order.assign(out_order, out_order + n_oo);
  l.mask.assign(mask, mask + n_mask);
  l.val.assign(val, val + n_val);
  l.col_id.assign(col_id, col_id + n_val);
  l.sparse_part = spMtxCRS<T>((size_t)m, (size_t)n, (size_t)sp_nz, sp_rst, sp_col, sp_val);
  std::vector<T> vb(b, b + n), vy(y, y + m);
  sparse_mv<T, spMtxLAV, false>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, l, sparse_matrix_descr{}, vb,
                                beta, vy);
  std::memcpy(y, vy.data(), sizeof(T) * m);
  return 0;
}

extern "C" {
int ref_cvr_spmv_f64(int64_t m, int64_t n, int64_t nitems, int simd_len, int nthreads,
                     const double* vals, const uint32_t* cols, int64_t nrec2, const int* rec,
                     const int* wb, const double* b, double alpha, double beta, double* y) {
  return cvr_run<double>(m, n, nitems, simd_len, nthreads, vals, cols, nrec2, rec, wb, b, alpha,
                         beta, y);
}
int ref_cvr_spmv_f32(int64_t m, int64_t n, int64_t nitems, int simd_len, int nthreads,
                     const float* vals, const uint32_t* cols, int64_t nrec2, const int* rec,
                     const int* wb, const float* b, float alpha, float beta, float* y) {
  return cvr_run<float>(m, n, nitems, simd_len, nthreads, vals, cols, nrec2, rec, wb, b, alpha,
                        beta, y);
}
int ref_lav_spmv_f64(int64_t m, int64_t n, int64_t nz, int lanes, int64_t sc, int64_t cc,
                     const int* co, int64_t n_co, const int* cf, int64_t n_cf, const uint64_t* sp,
                     int64_t n_sp, const uint32_t* oo, int64_t n_oo, const uint16_t* mask,
                     int64_t n_mask, const double* val, const uint32_t* col_id, int64_t n_val,
                     int64_t sp_nz, const int* sp_rst, const int* sp_col, const double* sp_val,
                     const double* b, double alpha, double beta, double* y) {
  return lav_run<double>(m, n, nz, lanes, sc, cc, co, n_co, cf, n_cf, sp, n_sp, oo, n_oo, mask,
                         n_mask, val, col_id, n_val, sp_nz, sp_rst, sp_col, sp_val, b, alpha, beta,
                         y);
}
int ref_lav_spmv_f32(int64_t m, int64_t n, int64_t nz, int lanes, int64_t sc, int64_t cc,
                     const int* co, int64_t n_co, const int* cf, int64_t n_cf, const uint64_t* sp,
                     int64_t n_sp, const uint32_t* oo, int64_t n_oo, const uint16_t* mask,
                     int64_t n_mask, const float* val, const uint32_t* col_id, int64_t n_val,
                     int64_t sp_nz, const int* sp_rst, const int* sp_col, const float* sp_val,
                     const float* b, float alpha, float beta, float* y) {
  return lav_run<float>(m, n, nz, lanes, sc, cc, co, n_co, cf, n_cf, sp, n_sp, oo, n_oo, mask,
                        n_mask, val, col_id, n_val, sp_nz, sp_rst, sp_col, sp_val, b, alpha, beta,
                        y);
}
int ref_sell_spmv_f64(int64_t m, int64_t n, int64_t nb, int C, int sigma, const int* Cl,
                      const int* Cs, const int* Bs, const int* SBs, int64_t slots,
                      const uint32_t* Col, const uint32_t* Perm, const double* Val,
                      const double* b, double alpha, double beta, double* y) {
  return sell_run<double>(m, n, nb, C, sigma, Cl, Cs, Bs, SBs, slots, Col, Perm, Val, b, alpha,
                          beta, y);
}
int ref_sell_spmv_f32(int64_t m, int64_t n, int64_t nb, int C, int sigma, const int* Cl,
                      const int* Cs, const int* Bs, const int* SBs, int64_t slots,
                      const uint32_t* Col, const uint32_t* Perm, const float* Val,
                      const float* b, float alpha, float beta, float* y) {
  return sell_run<float>(m, n, nb, C, sigma, Cl, Cs, Bs, SBs, slots, Col, Perm, Val, b, alpha,
                         beta, y);
}

}  // extern "C"
