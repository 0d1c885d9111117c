// api_drop_in.cpp -- uses the library exactly as code written against the
// reference API would: spMtxCRS, convert<T, Fmt, true>, sparse_mv<T, Fmt,
// true, stage> with std::vector operands, for all eight formats.  Prints one
// line per (format, type, stage form) with the largest row error against a
// plain CSR loop, scaled by the row's sum |a||x| + |beta||y|, and exits 1 if
// any exceeds 4*(len+2)*eps.  Run on a machine with a HIP device.
//
//   api_drop_in [G] [dump-dir]
//
// G: the Laplacian grid (default 100).  With a dump directory every product
// is also written there, one file each (operands, result and the matrix, see
// dump()), for a checker outside the program to compare with its own
// reference.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "rvvla_hip.h"
#include "sparse_matrix.h"

using namespace SparseMatrixLib;

template <class T>
spMtxCRS<T> laplacian(int g) {
  const int n = g * g, nnz = 5 * n - 4 * g;
  spMtxCRS<T> a(n, n, nnz);
  int k = 0;
  for (int r = 0; r < n; ++r) {
    a.Rst[r] = k;
    const int gy = r / g, gx = r % g;
    auto put = [&](int c, T v) { a.Col[k] = c; a.Val[k++] = v; };
    if (gy > 0) put(r - g, -1);
    if (gx > 0) put(r - 1, -1);
    put(r, 4);
    if (gx < g - 1) put(r + 1, -1);
    if (gy < g - 1) put(r + g, -1);
  }
  a.Rst[n] = k;
  return a;
}

// A ragged matrix: every 7th row of the Laplacian emptied, row 3 made
// 3000 entries long (columns spread over the grid).
template <class T>
spMtxCRS<T> ragged(int g) {
  const spMtxCRS<T> lap = laplacian<T>(g);
  const int n = (int)lap.n, extra = 3000;
  std::vector<int> rst(1, 0), col;
  std::vector<T> val;
  for (int r = 0; r < n; ++r) {
    if (r % 7 == 6) {
      rst.push_back((int)col.size());
      continue;
    }
    if (r == 3)
      for (int k = 0; k < extra; ++k) {
        col.push_back((int)((int64_t)k * n / extra));
        val.push_back((T)(1.0 + 0.001 * k));
      }
    else
      for (int e = lap.Rst[r]; e < lap.Rst[r + 1]; ++e) {
        col.push_back(lap.Col[e]);
        val.push_back(lap.Val[e]);
      }
    rst.push_back((int)col.size());
  }
  return spMtxCRS<T>((size_t)n, (size_t)n, col.size(), rst.data(), col.data(), val.data());
}

static int failures = 0;
static const char* g_dump = nullptr;
static int g_dumped = 0;

// One product as a file: int64 m, n, nnz, element bytes, transposed (0/1);
// double alpha, beta; int32 Rst[m+1], Col[nnz]; T Val[nnz]; T b, y0, y
// (b of n entries and y0, y of m, or the other way round when transposed).
template <class T>
void dump(const char* fmt, const char* form, const spMtxCRS<T>& a, const std::vector<T>& b,
          const std::vector<T>& y0, T alpha, T beta, const std::vector<T>& y, bool transposed) {
  if (!g_dump) return;
  char path[1024];
  std::snprintf(path, sizeof path, "%s/%03d_%s_%s_%s.bin", g_dump, g_dumped++, fmt,
                std::is_same<T, double>::value ? "double" : "float", form);
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    std::printf("cannot write %s\n", path);
    ++failures;
    return;
  }
  const int64_t hdr[5] = {(int64_t)a.m, (int64_t)a.n, (int64_t)a.nz, (int64_t)sizeof(T), transposed ? 1 : 0};
  const double ab[2] = {(double)alpha, (double)beta};
  std::fwrite(hdr, sizeof hdr, 1, f);
  std::fwrite(ab, sizeof ab, 1, f);
  std::fwrite(a.Rst, sizeof(int), a.m + 1, f);
  std::fwrite(a.Col, sizeof(int), a.nz, f);
  std::fwrite(a.Val, sizeof(T), a.nz, f);
  std::fwrite(b.data(), sizeof(T), b.size(), f);
  std::fwrite(y0.data(), sizeof(T), y0.size(), f);
  std::fwrite(y.data(), sizeof(T), y.size(), f);
  std::fclose(f);
}

static void report(const char* fmt, bool dbl, const char* what, bool same) {
  if (!same) ++failures;
  std::printf("%-5s %-6s %s %s\n", fmt, dbl ? "double" : "float", what, same ? "ok" : "FAIL");
}

template <class V>
static bool same_bytes(const std::vector<V>& a, const std::vector<V>& b) {
  return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(V)) == 0);
}

template <class T>
void check(const char* fmt, const char* form, const spMtxCRS<T>& a, const std::vector<T>& b,
           const std::vector<T>& y0, T alpha, T beta, const std::vector<T>& y,
           sparse_matrix_status st) {
  const double eps = std::is_same<T, double>::value ? 1.1102230246251565e-16 : 5.960464477539063e-08;
  double worst = 0;
  for (size_t i = 0; i < a.m; ++i) {
    double s = 0, sa = 0;
    for (int e = a.Rst[i]; e < a.Rst[i + 1]; ++e) {
      s += (double)a.Val[e] * b[a.Col[e]];
      sa += std::fabs((double)a.Val[e] * b[a.Col[e]]);
    }
    const double ref = alpha * s + beta * (double)y0[i];
    const double scale = std::fabs(alpha) * sa + std::fabs(beta) * std::fabs((double)y0[i]);
    const double len = a.Rst[i + 1] - a.Rst[i];
    const double err = std::fabs((double)y[i] - ref) / (4 * (len + 2) * eps * (scale > 0 ? scale : 1));
    if (err > worst) worst = err;
  }
  const bool ok = st.code == 0 && worst <= 1.0;
  if (!ok) ++failures;
  if (st.code == 0) dump<T>(fmt, form, a, b, y0, alpha, beta, y, false);
  std::printf("%-5s %-6s %-7s status=%d err/bound=%.3f %s\n", fmt,
              std::is_same<T, double>::value ? "double" : "float", form, st.code, worst,
              ok ? "ok" : "FAIL");
}

// y = alpha * A^T b + beta * y0 against a plain transposed loop
template <class T>
void check_transpose(const char* fmt, const spMtxCRS<T>& a, const std::vector<T>& b,
                     const std::vector<T>& y0, T alpha, T beta, const std::vector<T>& y,
                     sparse_matrix_status st) {
  const size_t n = a.n;
  std::vector<double> ref(n, 0.0), sa(n, 0.0);
  std::vector<int> len(n, 0);
  for (size_t r = 0; r < a.m; ++r)
    for (int e = a.Rst[r]; e < a.Rst[r + 1]; ++e) {
      ref[a.Col[e]] += (double)a.Val[e] * b[r];
      sa[a.Col[e]] += std::fabs((double)a.Val[e] * b[r]);
      len[a.Col[e]]++;
    }
  const double eps = std::is_same<T, double>::value ? 1.1102230246251565e-16 : 5.960464477539063e-08;
  double worst = 0;
  for (size_t c = 0; c < n; ++c) {
    const double want = alpha * ref[c] + beta * (double)y0[c];
    const double scale = std::fabs(alpha) * sa[c] + std::fabs(beta) * std::fabs((double)y0[c]);
    const double err = std::fabs((double)y[c] - want) / (4 * (len[c] + 2) * eps * (scale > 0 ? scale : 1));
    if (err > worst) worst = err;
  }
  const bool ok = st.code == 0 && worst <= 1.0;
  if (!ok) ++failures;
  if (st.code == 0) dump<T>(fmt, "transposed", a, b, y0, alpha, beta, y, true);
  std::printf("%-5s %-6s A^T     status=%d err/bound=%.3f %s\n", fmt,
              std::is_same<T, double>::value ? "double" : "float", st.code, worst, ok ? "ok" : "FAIL");
}

template <class T, template <class> class Fmt>
void run(const char* name, const spMtxCRS<T>& a, const convertParams& P) {
  std::vector<T> b(a.n), y0(a.m);
  for (size_t i = 0; i < a.n; ++i) b[i] = (T)std::sin(0.37 * (double)i);
  for (size_t i = 0; i < a.m; ++i) y0[i] = (T)std::cos(0.11 * (double)i);
  const T alpha = (T)1.5, beta = (T)-0.5;
  Fmt<T> d;
  convert<T, Fmt, true>(d, a, P);
  std::vector<T> y = y0;
  auto st = sparse_mv<T, Fmt, true>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, d,
                                    sparse_matrix_descr{}, b, beta, y);
  check<T>(name, "all", a, b, y0, alpha, beta, y, st);
  y = y0;
  auto s1 = sparse_mv<T, Fmt, true, SPARSE_MATRIX_MV_PREPARATION>(SPARSE_OPERATION_NON_TRANSPOSE,
                                                                  alpha, d, sparse_matrix_descr{}, b, beta, y);
  auto s2 = sparse_mv<T, Fmt, true, SPARSE_MATRIX_MV_MULT>(SPARSE_OPERATION_NON_TRANSPOSE, alpha, d,
                                                           sparse_matrix_descr{}, b, beta, y);
  auto s3 = sparse_mv<T, Fmt, true, SPARSE_MATRIX_MV_GET_RESULTS>(SPARSE_OPERATION_NON_TRANSPOSE,
                                                                  alpha, d, sparse_matrix_descr{}, b, beta, y);
  sparse_matrix_status s;
  s.code = s1.code ? s1.code : (s2.code ? s2.code : s3.code);
  check<T>(name, "staged", a, b, y0, alpha, beta, y, s);
  // the transposed operation: A^T b from the CRS the convert kept
  std::vector<T> bt(a.m), yt0(a.n);
  for (size_t i = 0; i < a.m; ++i) bt[i] = (T)std::sin(0.29 * (double)i);
  for (size_t i = 0; i < a.n; ++i) yt0[i] = (T)std::cos(0.17 * (double)i);
  std::vector<T> yt = yt0;
  auto st_t = sparse_mv<T, Fmt, true>(SPARSE_OPERATION_TRANSPOSE, alpha, d, sparse_matrix_descr{}, bt,
                                      beta, yt);
  check_transpose<T>(name, a, bt, yt0, alpha, beta, yt, st_t);
}

// SPARSE_OPERATION_TRANSPOSE on CRS: y = alpha * A^T b + beta * y, for a
// non-symmetric matrix (row-scaled Laplacian plus one extra column),
// against a plain transposed loop, in the all-stages and staged forms.
template <class T>
void transpose_crs(const spMtxCRS<T>& lap) {
  const size_t m = lap.m, n = lap.n + 3;
  spMtxCRS<T> a(m, n, lap.nz + m);
  size_t k = 0;
  for (size_t r = 0; r < m; ++r) {
    a.Rst[r] = (int)k;
    for (int e = lap.Rst[r]; e < lap.Rst[r + 1]; ++e) {
      a.Col[k] = lap.Col[e];
      a.Val[k++] = lap.Val[e] * (T)(1.0 + 0.001 * (double)r);
    }
    a.Col[k] = (int)(lap.n + r % 3);
    a.Val[k++] = (T)(0.5 + 0.25 * (double)(r % 5));
  }
  a.Rst[m] = (int)k;
  std::vector<T> b(m), y0(n);
  for (size_t i = 0; i < m; ++i) b[i] = (T)std::sin(0.21 * (double)i);
  for (size_t i = 0; i < n; ++i) y0[i] = (T)std::cos(0.13 * (double)i);
  const T alpha = (T)0.75, beta = (T)2;
  std::vector<double> ref(n, 0.0), sa(n, 0.0);
  std::vector<int> len(n, 0);
  for (size_t r = 0; r < m; ++r)
    for (int e = a.Rst[r]; e < a.Rst[r + 1]; ++e) {
      ref[a.Col[e]] += (double)a.Val[e] * b[r];
      sa[a.Col[e]] += std::fabs((double)a.Val[e] * b[r]);
      len[a.Col[e]]++;
    }
  const double eps = std::is_same<T, double>::value ? 1.1102230246251565e-16 : 5.960464477539063e-08;
  for (int form = 0; form < 2; ++form) {
    std::vector<T> y = y0;
    sparse_matrix_status st;
    if (form == 0) {
      st = sparse_mv<T, spMtxCRS, true>(SPARSE_OPERATION_TRANSPOSE, alpha, a, sparse_matrix_descr{},
                                        b, beta, y);
    } else {
      auto s1 = sparse_mv<T, spMtxCRS, true, SPARSE_MATRIX_MV_PREPARATION>(
          SPARSE_OPERATION_TRANSPOSE, alpha, a, sparse_matrix_descr{}, b, beta, y);
      auto s2 = sparse_mv<T, spMtxCRS, true, SPARSE_MATRIX_MV_MULT>(
          SPARSE_OPERATION_TRANSPOSE, alpha, a, sparse_matrix_descr{}, b, beta, y);
      auto s3 = sparse_mv<T, spMtxCRS, true, SPARSE_MATRIX_MV_GET_RESULTS>(
          SPARSE_OPERATION_TRANSPOSE, alpha, a, sparse_matrix_descr{}, b, beta, y);
      st.code = s1.code ? s1.code : (s2.code ? s2.code : s3.code);
    }
    double worst = 0;
    for (size_t c = 0; c < n; ++c) {
      const double want = alpha * ref[c] + beta * (double)y0[c];
      const double scale = std::fabs(alpha) * sa[c] + std::fabs(beta) * std::fabs((double)y0[c]);
      const double err = std::fabs((double)y[c] - want) / (4 * (len[c] + 2) * eps * (scale > 0 ? scale : 1));
      if (err > worst) worst = err;
    }
    const bool ok = st.code == 0 && worst <= 1.0;
    if (!ok) ++failures;
    std::printf("csr^T %-6s %-7s status=%d err/bound=%.3f %s\n",
                std::is_same<T, double>::value ? "double" : "float", form ? "staged" : "all", st.code,
                worst, ok ? "ok" : "FAIL");
  }
}

// CSR5, LAV, VNEC, VHCC and CVR built on the device ("device" = 1): the
// same fields as the host build, and the products of the device-built
// containers
template <class T>
void device_built(const spMtxCRS<T>& a) {
  const bool dbl = std::is_same<T, double>::value;
  convertParams none, dev;
  dev.param["device"].i = 1;
  {
    spMtxCSR5<T> h, d;
    convert<T, spMtxCSR5, true>(h, a, none);
    convert<T, spMtxCSR5, true>(d, a, dev);
    report("csr5", dbl, "device-built fields",
           h.format() == d.format() && h._p == d._p && h._sigma == d._sigma &&
               h._bit_y_offset == d._bit_y_offset && h._bit_seg_offset == d._bit_seg_offset &&
               h._tile_desc_len == d._tile_desc_len && h._tail_tile_start == d._tail_tile_start &&
               h._num_offsets == d._num_offsets && h._csr_row_pointer == d._csr_row_pointer &&
               h._csr_column_index == d._csr_column_index && same_bytes(h._csr_value, d._csr_value) &&
               h._csr5_tile_pointer == d._csr5_tile_pointer &&
               h._csr5_tile_descriptor == d._csr5_tile_descriptor &&
               h._csr5_tile_empty_offset_pointer == d._csr5_tile_empty_offset_pointer &&
               h._csr5_tile_empty_offset == d._csr5_tile_empty_offset);
    run<T, spMtxCSR5>("csr5", a, dev);
  }
  {
    spMtxLAV<T> h, d;
    convert<T, spMtxLAV, true>(h, a, none);
    convert<T, spMtxLAV, true>(d, a, dev);
    const auto& hs = h.sparse_part;
    const auto& ds = d.sparse_part;
    const bool sp = hs.m == ds.m && hs.nz == ds.nz &&
                    std::memcmp(hs.Rst, ds.Rst, (hs.m + 1) * sizeof(int)) == 0 &&
                    (hs.nz == 0 || (std::memcmp(hs.Col, ds.Col, hs.nz * sizeof(int)) == 0 &&
                                    std::memcmp(hs.Val, ds.Val, hs.nz * sizeof(T)) == 0));
    report("lav", dbl, "device-built fields",
           sp && h.segment_count == d.segment_count && h.chunk_count == d.chunk_count &&
               h.chunk_offsets == d.chunk_offsets &&
               h.chunk_offsets_full_mask == d.chunk_offsets_full_mask &&
               h.segment_ptr == d.segment_ptr && h.out_order == d.out_order && h.mask == d.mask &&
               same_bytes(h.val, d.val) && h.col_id == d.col_id);
    run<T, spMtxLAV>("lav", a, dev);
  }
  {
    spMtxVNEC<T> h, d;
    convert<T, spMtxVNEC, true>(h, a, none);
    convert<T, spMtxVNEC, true>(d, a, dev);
    report("vnec", dbl, "device-built fields",
           h.NUM_THREADS_VNEC == d.NUM_THREADS_VNEC &&
               same_bytes(h.thread_coord_start, d.thread_coord_start) &&
               same_bytes(h.thread_coord_end, d.thread_coord_end) && h.ecr_indices == d.ecr_indices &&
               h.NEC_NUM == d.NEC_NUM && h.nec_ptr == d.nec_ptr &&
               h.use_x_indices == d.use_x_indices && h.IRD_mat == d.IRD_mat && h.vnec_t == d.vnec_t);
    run<T, spMtxVNEC>("vnec", a, dev);
  }
  {
    spMtxVHCC<T> h, d;
    convert<T, spMtxVHCC, true>(h, a, none);
    convert<T, spMtxVHCC, true>(d, a, dev);
    const auto hp = h.get_params(), dp = d.get_params();
    report("vhcc", dbl, "device-built fields",
           hp.num_threads == dp.num_threads && hp.vec_size == dp.vec_size &&
               hp.n_masks == dp.n_masks && h.get_num_vectors() == d.get_num_vectors() &&
               h.get_pad_entries() == d.get_pad_entries() && h.get_pad_rows() == d.get_pad_rows() &&
               same_bytes(h.thr_info(), d.thr_info()) && h.veceor() == d.veceor() &&
               h.scan_mask() == d.scan_mask() && h.row_arr() == d.row_arr() &&
               h.col_arr() == d.col_arr() && same_bytes(h.vals_arr(), d.vals_arr()));
    run<T, spMtxVHCC>("vhcc", a, dev);
  }
  {
    spMtxCVR<T> h, d;
    convert<T, spMtxCVR, true>(h, a, none);
    convert<T, spMtxCVR, true>(d, a, dev);
    report("cvr", dbl, "device-built fields",
           h.nItems == d.nItems && h.Nthreads == d.Nthreads && h.SIMD_LEN == d.SIMD_LEN &&
               same_bytes(h.vPack_vec_vals, d.vPack_vec_vals) &&
               h.vPack_vec_cols == d.vPack_vec_cols && h.vPack_vec_record == d.vPack_vec_record &&
               h.writeback_pos_record == d.writeback_pos_record);
    run<T, spMtxCVR>("cvr", a, dev);
  }
}

template <class T>
void all_formats(int g) {
  const spMtxCRS<T> a = laplacian<T>(g);
  convertParams none, hp;
  hp.param["R"].i = 300;
  hp.param["C"].i = 500;
  hp.param["p"].d = 2.0;
  // CRS itself multiplies without a convert call (device image on first use)
  {
    std::vector<T> b(a.n, (T)1), y(a.m, (T)0), y0 = y;
    auto st = sparse_mv<T, spMtxCRS, true>(SPARSE_OPERATION_NON_TRANSPOSE, (T)1, a,
                                           sparse_matrix_descr{}, b, (T)0, y);
    check<T>("csr", "all", a, b, y0, (T)1, (T)0, y, st);
  }
  run<T, spMtxHCSR>("hcsr", a, hp);
  // HCSR built on the device ("device" = 1): same fields as the host build
  {
    convertParams dev = hp;
    dev.param["device"].i = 1;
    spMtxHCSR<T> h, d;
    convert<T, spMtxHCSR, true>(h, a, hp);
    convert<T, spMtxHCSR, true>(d, a, dev);
    const bool same = h.R == d.R && h.C == d.C && h.b_m == d.b_m && h.b_n == d.b_n &&
                      h.block_ptr == d.block_ptr && h.row_offset == d.row_offset &&
                      h.row_ptr == d.row_ptr && h.col_idx == d.col_idx && h.values == d.values;
    if (!same) ++failures;
    std::printf("hcsr  %-6s device-built fields %s\n", std::is_same<T, double>::value ? "double" : "float",
                same ? "ok" : "FAIL");
    run<T, spMtxHCSR>("hcsr", a, dev);
  }
  // SELL built on the device ("device" = 1): same fields as the host build
  {
    convertParams dev;
    dev.param["device"].i = 1;
    spMtxSELL_C_Sigma<T> h, d;
    convert<T, spMtxSELL_C_Sigma, true>(h, a, none);
    convert<T, spMtxSELL_C_Sigma, true>(d, a, dev);
    const bool same = h.Sigma == d.Sigma && h.cnt_b == d.cnt_b && h.vCl == d.vCl &&
                      h.vCs == d.vCs && h.vBs == d.vBs && h.vSBs == d.vSBs && h.vCol == d.vCol &&
                      h.vPerm == d.vPerm && h.vVal == d.vVal;
    if (!same) ++failures;
    std::printf("sell  %-6s device-built fields %s\n", std::is_same<T, double>::value ? "double" : "float",
                same ? "ok" : "FAIL");
    convertParams dev16 = dev;
    dev16.param["sigma"].i = 16;
    run<T, spMtxSELL_C_Sigma>("sell", a, dev16);
  }
  run<T, spMtxSELL_C_Sigma>("sell", a, none);
  device_built<T>(a);
  device_built<T>(ragged<T>(g));
  run<T, spMtxCSR5>("csr5", a, none);
  run<T, spMtxLAV>("lav", a, none);
  run<T, spMtxVNEC>("vnec", a, none);
  run<T, spMtxVHCC>("vhcc", a, none);
  run<T, spMtxCVR>("cvr", a, none);
  transpose_crs<T>(a);
}

int main(int argc, char** argv) {
  const int g = argc > 1 ? std::atoi(argv[1]) : 100;
  if (argc > 2) g_dump = argv[2];
  all_formats<double>(g);
  all_formats<float>(g);
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
