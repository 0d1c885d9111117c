// rvvla_accuracy -- accuracy check of every format and precision through
// the C-ABI, on a matrix file or a generated matrix.
//
//   rvvla_accuracy <matrix> [--formats all|csr,sell,...] [--dtypes f64,f32,bf16,f16]
//                  [--alpha A] [--beta B]
//
// <matrix>: a Matrix Market (.mtx) or binary CSR (.bin) file, lap:G,
// powerlaw:N[:NNZ] or random:N:K.  For each (format, dtype) the matrix is
// converted, uploaded and multiplied once (y = alpha*A*x + beta*y, x and y
// uniform in [-1, 1]); every row is compared with a long-double row sum on
// the host.  Bound per row: FP64/FP32, 4*(len+2)*u*(|alpha|*sum|a||x| +
// |beta||y0|) with the values rounded to the dtype; BF16/FP16 values,
// (u_v + (4*len+10)*2^-24) * the same scale against the unrounded values
// (u_v = 2^-8 / 2^-11).  Prints the worst error/bound and the normwise
// relative error per line; exit status 1 if any row exceeds its bound
// (a format that refuses the matrix is reported, not counted).
#include "common.h"

using namespace apps;


int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <matrix> [--formats all] [--dtypes f64,f32,bf16,f16] "
                         "[--alpha A] [--beta B]\n", argv[0]);
    return 2;
  }
  std::string fmts = "all", dts = "f64,f32,bf16,f16";
  double alpha = 1.5, beta = -0.5;
  for (int i = 2; i + 1 < argc; i += 2) {
    const std::string k = argv[i];
    if (k == "--formats") fmts = argv[i + 1];
    else if (k == "--dtypes") dts = argv[i + 1];
    else if (k == "--alpha") alpha = std::atof(argv[i + 1]);
    else if (k == "--beta") beta = std::atof(argv[i + 1]);
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  int failures = 0;
  try {
    Matrix M;
    load(argv[1], M);
    const rvvla_gen_csr& g = M.g;
    std::printf("matrix %s: %lld x %lld, %lld entries\n", M.name.c_str(), (long long)g.m,
                (long long)g.n, (long long)g.nnz);
    std::vector<double> x64((size_t)g.n), y064((size_t)g.m);
    ck(rvvla_gen_uniform(g.n, -1, 1, 7, 1, x64.data()), "x");
    ck(rvvla_gen_uniform(g.m, -1, 1, 7, 2, y064.data()), "y");
    for (int dtype : parse_dtypes(dts)) {
      const bool narrow = dtype == RVVLA_BF16 || dtype == RVVLA_F16;
      const double u = dtype == RVVLA_F64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
      const double uv = dtype == RVVLA_BF16 ? std::ldexp(1.0, -8) : std::ldexp(1.0, -11);
      // operands as the device sees them
      std::vector<double> xs((size_t)g.n), y0s((size_t)g.m), vs((size_t)g.nnz);
      for (int64_t i = 0; i < g.n; ++i) xs[i] = dtype == RVVLA_F64 ? x64[i] : (double)(float)x64[i];
      for (int64_t i = 0; i < g.m; ++i) y0s[i] = dtype == RVVLA_F64 ? y064[i] : (double)(float)y064[i];
      for (int64_t e = 0; e < g.nnz; ++e)
        vs[e] = dtype == RVVLA_F64 ? g.val[e] : narrow ? g.val[e] : (double)(float)g.val[e];
      // reference rows (long double) and scales
      std::vector<long double> ref((size_t)g.m);
      std::vector<double> scale((size_t)g.m);
      for (int64_t r = 0; r < g.m; ++r) {
        long double s = 0;
        double sa = 0;
        for (int32_t e = g.rst[r]; e < g.rst[r + 1]; ++e) {
          s += (long double)vs[e] * xs[g.col[e]];
          sa += std::fabs(vs[e] * xs[g.col[e]]);
        }
        ref[r] = (long double)alpha * s + (long double)beta * y0s[r];
        scale[r] = std::fabs(alpha) * sa + std::fabs(beta) * std::fabs(y0s[r]);
      }
      const Values V = values_for(g, dtype);
      const size_t vb = vec_bytes(dtype);
      std::vector<unsigned char> xh((size_t)g.n * vb), yh((size_t)g.m * vb);
      for (int64_t i = 0; i < g.n; ++i) {
        if (vb == 8) std::memcpy(&xh[i * 8], &xs[i], 8);
        else { const float f = (float)xs[i]; std::memcpy(&xh[i * 4], &f, 4); }
      }
      DevBuf dx(xh.size()), dy(yh.size());
      hk(hipMemcpy(dx.p, xh.data(), xh.size(), hipMemcpyHostToDevice), "copy x");
      for (int f : parse_formats(fmts)) {
        rvvla_layout_t L = nullptr;
        rvvla_mat_t A = nullptr;
        int rc = rvvla_convert(f, dtype, g.m, g.n, g.nnz, g.rst, g.col, V.ptr(dtype), "", &L);
        if (!rc) rc = rvvla_upload(L, &A);
        if (rc) {
          std::printf("%-5s %-5s refused: %s\n", rvvla_format_name(f), dtype_name(dtype),
                      rvvla_strerror(rc));
          if (L) rvvla_layout_free(L);
          continue;
        }
        for (int64_t i = 0; i < g.m; ++i) {
          if (vb == 8) std::memcpy(&yh[i * 8], &y0s[i], 8);
          else { const float f32 = (float)y0s[i]; std::memcpy(&yh[i * 4], &f32, 4); }
        }
        hk(hipMemcpy(dy.p, yh.data(), yh.size(), hipMemcpyHostToDevice), "copy y");
        ck(rvvla_spmv(A, alpha, dx.p, beta, dy.p, nullptr), "spmv");
        hk(hipDeviceSynchronize(), "sync");
        hk(hipMemcpy(yh.data(), dy.p, yh.size(), hipMemcpyDeviceToHost), "copy y back");
        double worst = 0, num = 0, den = 0;
        for (int64_t r = 0; r < g.m; ++r) {
          double got;
          if (vb == 8) std::memcpy(&got, &yh[r * 8], 8);
          else { float f32; std::memcpy(&f32, &yh[r * 4], 4); got = f32; }
          const double len = g.rst[r + 1] - g.rst[r];
          const double bound = narrow ? (uv + (4 * len + 10) * std::ldexp(1.0, -24)) * scale[r]
                                      : 4 * (len + 2) * u * scale[r];
          const double err = std::fabs(got - (double)ref[r]);
          worst = std::max(worst, bound > 0 ? err / bound : (err > 0 ? INFINITY : 0.0));
          num += err * err;
          den += (double)(ref[r] * ref[r]);
        }
        const bool ok = worst <= 1.0;
        failures += !ok;
        std::printf("%-5s %-5s err/bound %.3f  normwise %.3e  %s\n", rvvla_format_name(f),
                    dtype_name(dtype), worst, den > 0 ? std::sqrt(num / den) : std::sqrt(num),
                    ok ? "ok" : "FAIL");
        rvvla_mat_free(A);
        rvvla_layout_free(L);
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
