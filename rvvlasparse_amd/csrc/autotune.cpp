// autotune.cpp -- picks the storage format of a CSR matrix and its
// parameters by timing candidate images on the device (no reference
// counterpart: the reference leaves the choice of format, SELL C and
// sigma, CSR5 omega and sigma and HCSR R, C and p to the caller).
//
// Matrices are keyed by their row-length histogram (row_length_key: the
// fractions of rows in power-of-two length buckets, rounded to 1/20, the
// bucket of max/mean) together with the power-of-two buckets of m and n,
// the value type and the candidate list; the winner is cached for the
// process and, with cache=<path>, in a text file of "key format params"
// lines.  Each candidate is converted and uploaded, run twice untimed and
// then `reps` times between two events on a stream of its own; candidates
// a format refuses (a layout it cannot hold) are skipped.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "devmat.h"
#include "internal.h"

namespace rvvla {

std::string row_length_key(int64_t m, const int* rst, int k) {
  int hist[32] = {0};
  int64_t mx = 0, tot = 0;
  for (int64_t r = 0; r < m; ++r) {
    const int64_t l = rst[r + 1] - rst[r];
    int b = 0;
    while ((int64_t(1) << b) <= l && b < 31) ++b;
    hist[b]++;
    mx = std::max(mx, l);
    tot += l;
  }
  std::ostringstream os;
  int kp = 1;
  while (kp < k) kp <<= 1;
  os << "k" << kp << ":";
  for (int b = 0; b < 32; ++b) {
    const int q = m ? (int)std::lround(20.0 * hist[b] / (double)m) : 0;
    if (q) os << b << "=" << q << ",";
  }
  const double ratio = tot ? (double)mx * (double)m / (double)tot : 1.0;
  os << "r" << (int)std::ceil(std::log2(std::max(1.0, ratio)));
  return os.str();
}

namespace {

struct Cand {
  int format;
  std::string params;
};

std::mutex g_mu;
std::map<std::string, std::pair<int, std::string>> g_cache;

int log2_bucket(int64_t v) {
  int b = 0;
  while ((int64_t(1) << b) < v && b < 62) ++b;
  return b;
}

// The candidates of one format for an m x n matrix.
void candidates(int format, int dtype, int64_t m, int64_t n, std::vector<Cand>& out) {
  switch (format) {
    case RVVLA_FMT_CSR:
      out.push_back({format, ""});  // the image chosen from the matrix
      out.push_back({format, "sell=0,xslices=0,xlds=0"});  // the plain adaptive kernel
      break;
    case RVVLA_FMT_SELL:
      // C is the wavefront (64); sigma: natural order, short windows, long windows
      for (const char* s : {"sigma=1", "sigma=64", "sigma=4096", ""}) out.push_back({format, s});
      break;
    case RVVLA_FMT_CSR5:  // omega is the wavefront (64); sigma per tile
      for (const char* s : {"sigma=4", "sigma=8", "sigma=16"}) out.push_back({format, s});
      break;
    case RVVLA_FMT_HCSR: {
      // the reference defaults (R = C = 2000 FP64, 4000 otherwise), square
      // power-of-two blocks and wide blocks; grids of more than 2^26 blocks
      // are left out (their block tables alone would dwarf the matrix)
      const int64_t geo[][2] = {{0, 0}, {2048, 2048}, {4096, 4096}, {256, 8192}, {1024, 16384}};
      for (const auto& g : geo) {
        if (g[0] == 0) {
          out.push_back({format, ""});
          continue;
        }
        const int64_t blocks = ((m + g[0] - 1) / g[0]) * ((n + g[1] - 1) / g[1]);
        if (blocks > (int64_t(1) << 26)) continue;
        out.push_back({format, "R=" + std::to_string(g[0]) + ",C=" + std::to_string(g[1])});
      }
      break;
    }
    default:
      out.push_back({format, ""});
  }
  (void)dtype;
}

double time_candidate(const Cand& c, int dtype, int64_t m, int64_t n, int64_t nnz,
                      const int32_t* rst, const int32_t* col, const void* val, const void* x,
                      void* y, hipStream_t s, int reps) {
  rvvla_layout_t L = nullptr;
  if (rvvla_convert(c.format, dtype, m, n, nnz, rst, col, val, c.params.c_str(), &L) != RVVLA_OK)
    return -1;
  rvvla_mat_t A = nullptr;
  const int rc = rvvla_upload(L, &A);
  rvvla_layout_free(L);
  if (rc != RVVLA_OK) return -1;
  double ms = -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    bool ok = true;
    for (int w = 0; w < 2 && ok; ++w) ok = rvvla_spmv(A, 1.0, x, 0.0, y, s) == RVVLA_OK;
    ok = ok && hipEventRecord(e0, s) == hipSuccess;
    for (int r = 0; r < reps && ok; ++r) ok = rvvla_spmv(A, 1.0, x, 0.0, y, s) == RVVLA_OK;
    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0;
    if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = (double)t / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  rvvla_mat_free(A);
  return ms;
}

void load_cache(const std::string& path) {
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key, prm;
    int fmt = -1;
    if (!(is >> key >> fmt)) continue;
    is >> prm;
    if (prm == "-") prm.clear();
    if (fmt >= 0 && fmt < RVVLA_NUM_FORMATS) g_cache[key] = {fmt, prm};
  }
}

void save_cache(const std::string& path) {
  std::ofstream f(path);
  for (const auto& kv : g_cache)
    f << kv.first << " " << kv.second.first << " " << (kv.second.second.empty() ? "-" : kv.second.second)
      << "\n";
}

}  // namespace
}  // namespace rvvla

extern "C" int rvvla_autotune(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* rst,
                              const int32_t* col, const void* val, const char* params, int* format,
                              char* out, int len) {
  using namespace rvvla;
  try {
    if (!format || !out || len < 1) return RVVLA_ERR_INVALID_ARG;
    if (dtype < RVVLA_F64 || dtype > RVVLA_F16) return RVVLA_ERR_VALUE_TYPE;
    check_csr_public(m, n, nnz, rst, col);
    if (nnz > 0 && !val) throw Error(RVVLA_ERR_INVALID_ARG, "autotune: null values");
    // the cache path and the format list are not numbers: taken out
    // verbatim (each ends at ',' or ';') before the rest is parsed
    std::string rest = params ? params : "";
    auto take = [&rest](const char* key, const std::string& dflt) {
      const size_t p = rest.find(key);
      if (p == std::string::npos) return dflt;
      const size_t b = p + std::strlen(key), e = rest.find_first_of(",;", b);
      const std::string v = rest.substr(b, e == std::string::npos ? std::string::npos : e - b);
      rest.erase(p, e == std::string::npos ? std::string::npos : e - p + 1);
      return v;
    };
    const std::string cache = take("cache=", ""), fmts = take("formats=", "all");
    const SparseMatrixLib::convertParams P = parse_params(rest.c_str());
    const int reps = param_int(P, "reps", 20);
    if (reps < 1 || reps > 10000) throw Error(RVVLA_ERR_INVALID_ARG, "autotune: reps in [1, 10000]");
    std::vector<int> fl;
    if (fmts == "all" || fmts.empty()) {
      for (int f = 0; f < RVVLA_NUM_FORMATS; ++f) fl.push_back(f);
    } else {
      std::istringstream is(fmts);
      std::string name;
      while (std::getline(is, name, '+')) {
        int f = -1;
        for (int g = 0; g < RVVLA_NUM_FORMATS; ++g)
          if (name == rvvla_format_name(g)) f = g;
        if (f < 0) throw Error(RVVLA_ERR_UNKNOWN_FORMAT, "autotune: unknown format " + name);
        fl.push_back(f);
      }
    }
    std::ostringstream ks;
    ks << "d" << dtype << ":m" << log2_bucket(m) << ":n" << log2_bucket(n) << ":f";
    for (int f : fl) ks << f;
    ks << ":" << row_length_key(m, rst, 1);
    const std::string key = ks.str();
    std::lock_guard<std::mutex> lock(g_mu);
    if (!cache.empty()) load_cache(cache);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) {
      const size_t vb = dtype == RVVLA_F64 ? 8 : 4;
      void *x = nullptr, *y = nullptr;
      const int64_t len_x = std::max<int64_t>(n, 1), len_y = std::max<int64_t>(m, 1);
      if (hip_check(hipMalloc(&x, (size_t)len_x * vb), "hipMalloc") ||
          hip_check(hipMalloc(&y, (size_t)len_y * vb), "hipMalloc")) {
        if (x) (void)hipFree(x);
        throw Error(RVVLA_ERR_HIP, last_error());
      }
      hipStream_t s = nullptr;
      (void)hipMemset(x, 0, (size_t)len_x * vb);
      const bool have_stream = hipStreamCreate(&s) == hipSuccess;
      std::vector<Cand> cands;
      for (int f : fl) candidates(f, dtype, m, n, cands);
      double best = -1;
      std::pair<int, std::string> win{-1, ""};
      for (const Cand& c : cands) {
        if (!have_stream) break;
        const double t = time_candidate(c, dtype, m, n, nnz, rst, col, val, x, y, s, reps);
        if (t > 0 && (best < 0 || t < best)) {
          best = t;
          win = {c.format, c.params};
        }
      }
      if (s) (void)hipStreamDestroy(s);
      (void)hipFree(x);
      (void)hipFree(y);
      if (win.first < 0) throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "autotune: no candidate ran");
      it = g_cache.emplace(key, win).first;
    }
    if (!cache.empty()) save_cache(cache);
    *format = it->second.first;
    if ((int)it->second.second.size() + 1 > len) throw Error(RVVLA_ERR_LENGTH, "autotune: output too short");
    std::snprintf(out, (size_t)len, "%s", it->second.second.c_str());
    return RVVLA_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RVVLA_ERR_LOGIC;
  }
}
