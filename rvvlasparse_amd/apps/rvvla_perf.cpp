// rvvla_perf -- SpMV throughput of every format and precision through the
// C-ABI, on a matrix file or a generated matrix.
//
//   rvvla_perf <matrix> [--formats all|csr,sell,...] [--dtypes f64,f32,bf16,f16]
//              [--iters N] [--warmup W] [--params "key=value,..."]
//
// --params is handed to every conversion (convertParams, e.g. "xslices=0"
// or "R=256,C=8192").
// <matrix> as for rvvla_accuracy.  For each (format, dtype): convert and
// upload once (timed on the host), W untimed products, then N back-to-back
// products y = A*x timed with HIP events on one stream.  Reports the time
// per product, the compulsory-byte rate B / t with B = nnz*(s_val + 4) +
// (m+1)*4 + n*s_x + m*s_y, its fraction of the 8 TB/s HBM peak, GFLOP/s
// (2*nnz / t) and the device bytes of the format.
#include <chrono>

#include "common.h"

using namespace apps;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <matrix> [--formats all] [--dtypes f64,f32,bf16,f16] "
                         "[--iters N] [--warmup W] [--params key=value,...]\n", argv[0]);
    return 2;
  }
  std::string fmts = "all", dts = "f64,f32", params;
  int iters = 100, warmup = 10;
  for (int i = 2; i + 1 < argc; i += 2) {
    const std::string k = argv[i];
    if (k == "--formats") fmts = argv[i + 1];
    else if (k == "--dtypes") dts = argv[i + 1];
    else if (k == "--iters") iters = std::max(1, std::atoi(argv[i + 1]));
    else if (k == "--warmup") warmup = std::max(0, std::atoi(argv[i + 1]));
    else if (k == "--params") params = argv[i + 1];
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  try {
    Matrix M;
    load(argv[1], M);
    const rvvla_gen_csr& g = M.g;
    std::printf("matrix %s: %lld x %lld, %lld entries; %d products after %d warm-up\n",
                M.name.c_str(), (long long)g.m, (long long)g.n, (long long)g.nnz, iters, warmup);
    std::printf("%-5s %-5s %12s %10s %8s %9s %12s %10s\n", "fmt", "dtype", "us/product", "GB/s",
                "%peak", "GFLOP/s", "device MB", "upload s");
    std::vector<double> xv((size_t)g.n);
    ck(rvvla_gen_uniform(g.n, -1, 1, 7, 1, xv.data()), "x");
    hipStream_t s;
    hk(hipStreamCreate(&s), "stream");
    hipEvent_t e0, e1;
    hk(hipEventCreate(&e0), "event");
    hk(hipEventCreate(&e1), "event");
    for (int dtype : parse_dtypes(dts)) {
      const Values V = values_for(g, dtype);
      const size_t vb = vec_bytes(dtype);
      std::vector<unsigned char> xh((size_t)g.n * vb);
      for (int64_t i = 0; i < g.n; ++i) {
        if (vb == 8) std::memcpy(&xh[i * 8], &xv[i], 8);
        else { const float f = (float)xv[i]; std::memcpy(&xh[i * 4], &f, 4); }
      }
      DevBuf dx(xh.size()), dy((size_t)g.m * vb);
      hk(hipMemcpy(dx.p, xh.data(), xh.size(), hipMemcpyHostToDevice), "copy x");
      for (int f : parse_formats(fmts)) {
        rvvla_layout_t L = nullptr;
        rvvla_mat_t A = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        int rc = rvvla_convert(f, dtype, g.m, g.n, g.nnz, g.rst, g.col, V.ptr(dtype), params.c_str(), &L);
        if (!rc) rc = rvvla_upload(L, &A);
        const double up = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc) {
          std::printf("%-5s %-5s refused: %s\n", rvvla_format_name(f), dtype_name(dtype),
                      rvvla_strerror(rc));
          if (L) rvvla_layout_free(L);
          continue;
        }
        for (int w = 0; w < warmup; ++w) ck(rvvla_spmv(A, 1.0, dx.p, 0.0, dy.p, s), "spmv");
        hk(hipEventRecord(e0, s), "record");
        for (int it = 0; it < iters; ++it) ck(rvvla_spmv(A, 1.0, dx.p, 0.0, dy.p, s), "spmv");
        hk(hipEventRecord(e1, s), "record");
        hk(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hk(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        const double t = ms * 1e-3 / iters;
        int64_t dbytes = 0;
        rvvla_mat_info(A, nullptr, nullptr, nullptr, nullptr, nullptr, &dbytes);
        const double gbs = algo_bytes(g, dtype) / t / 1e9;
        std::printf("%-5s %-5s %12.2f %10.1f %7.1f%% %9.1f %12.1f %10.2f\n", rvvla_format_name(f),
                    dtype_name(dtype), t * 1e6, gbs, 100.0 * gbs / 8000.0,
                    2.0 * (double)g.nnz / t / 1e9, dbytes / 1e6, up);
        std::fflush(stdout);
        rvvla_mat_free(A);
        rvvla_layout_free(L);
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
