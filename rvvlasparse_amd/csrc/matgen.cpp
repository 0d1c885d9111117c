// matgen.cpp -- deterministic synthetic matrices for tests and benchmarks.
//
// Every random draw comes from a splitmix64 stream keyed by (seed, purpose,
// index), so a row's contents depend only on the seed and the row number,
// never on the thread count.  Purposes: 1 = row structure, 2 = values,
// 3+ = dense vectors (caller's choice).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "internal.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t purpose, uint64_t index)
      : s(mix64(mix64(seed + 0x9E3779B97F4A7C15ull * (purpose + 1)) ^ (index * 0xD1B54A32D192ED03ull))) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
  }
  // uniform in [0, 1) with 53 random bits
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
  // uniform integer in [0, n)
  int64_t below(int64_t n) { return (int64_t)(((unsigned __int128)next() * (uint64_t)n) >> 64); }
  double normal() {
    const double u1 = 1.0 - u01();  // (0, 1]
    const double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

// Append draws from [lo, hi] until `v` holds `need` distinct sorted values.
void top_up(std::vector<int>& v, size_t need, int64_t lo, int64_t hi, Rng& r) {
  while (v.size() < need) {
    const size_t missing = need - v.size();
    for (size_t k = 0; k < missing; ++k) v.push_back((int)(lo + r.below(hi - lo + 1)));
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
}

// `v` (empty on entry) receives k distinct sorted values of [lo, hi]: drawn
// directly when k is at most half the range, otherwise the range minus
// (range - k) drawn values.
void pick_distinct(std::vector<int>& v, size_t k, int64_t lo, int64_t hi, Rng& r) {
  const size_t w = (size_t)(hi - lo + 1);
  if (2 * k <= w) {
    top_up(v, k, lo, hi, r);
    return;
  }
  std::vector<int> skip;
  top_up(skip, w - k, lo, hi, r);
  size_t s = 0;
  for (int64_t c = lo; c <= hi; ++c) {
    if (s < skip.size() && skip[s] == c) {
      ++s;
      continue;
    }
    v.push_back((int)c);
  }
}

int alloc_out(rvvla_gen_csr* out, int64_t m, int64_t n, int64_t nnz) {
  out->m = m;
  out->n = n;
  out->nnz = nnz;
  out->rst = (int32_t*)std::malloc(sizeof(int32_t) * (m + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  out->val = (double*)std::malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
  if (!out->rst || !out->col || !out->val) {
    rvvla_gen_free(out);
    return RVVLA_ERR_OUT_OF_MEMORY;
  }
  return RVVLA_OK;
}

}  // namespace

extern "C" {

int rvvla_gen_free(rvvla_gen_csr* g) {
  if (!g) return RVVLA_ERR_INVALID_ARG;
  std::free(g->rst);
  std::free(g->col);
  std::free(g->val);
  g->rst = nullptr;
  g->col = nullptr;
  g->val = nullptr;
  return RVVLA_OK;
}

int rvvla_gen_laplacian5(int64_t g, rvvla_gen_csr* out) {
  if (!out || g < 1 || g > 40000) return RVVLA_ERR_INVALID_ARG;
  const int64_t n = g * g;
  const int64_t nnz = 5 * n - 4 * g;
  if (int rc = alloc_out(out, n, n, nnz)) return rc;
  // Row r = gy*g + gx; entries before row r: 5r minus the neighbours that
  // fall off the grid in rows < r.
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r <= n; ++r) {
    const int64_t gy = r / g, gx = r % g;
    // missing neighbours among rows [0, r): up (first grid row), down (last grid
    // row), left (first column), right (last column)
    const int64_t up = std::min<int64_t>(r, g);
    const int64_t down = std::max<int64_t>(0, r - (n - g));
    const int64_t left = gy + (gx > 0 ? 1 : 0);
    const int64_t right = gy;
    out->rst[r] = (int32_t)(5 * r - up - down - left - right);
  }
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    const int64_t gy = r / g, gx = r % g;
    int64_t k = out->rst[r];
    if (gy > 0) { out->col[k] = (int32_t)(r - g); out->val[k++] = -1.0; }
    if (gx > 0) { out->col[k] = (int32_t)(r - 1); out->val[k++] = -1.0; }
    out->col[k] = (int32_t)r; out->val[k++] = 4.0;
    if (gx < g - 1) { out->col[k] = (int32_t)(r + 1); out->val[k++] = -1.0; }
    if (gy < g - 1) { out->col[k] = (int32_t)(r + g); out->val[k++] = -1.0; }
  }
  return RVVLA_OK;
}

int rvvla_gen_random_rows(int64_t n, int k, int include_diag, int normal, double stdev,
                          uint64_t seed, rvvla_gen_csr* out) {
  if (!out || n < 1 || k < 1 || k > n || n > 2000000000) return RVVLA_ERR_INVALID_ARG;
  const int64_t nnz = n * (int64_t)k;
  if (nnz > 2147483647LL) return RVVLA_ERR_LENGTH;
  if (int rc = alloc_out(out, n, n, nnz)) return rc;
  for (int64_t r = 0; r <= n; ++r) out->rst[r] = (int32_t)(r * k);
#pragma omp parallel
  {
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
      Rng rs(seed, 1, (uint64_t)r), rv(seed, 2, (uint64_t)r);
      cols.clear();
      if (include_diag) cols.push_back((int)r);
      top_up(cols, (size_t)k, 0, n - 1, rs);
      int64_t base = r * k;
      for (int j = 0; j < k; ++j) {
        out->col[base + j] = cols[j];
        out->val[base + j] = normal ? stdev * rv.normal() : 2.0 * rv.u01() - 1.0;
      }
    }
  }
  return RVVLA_OK;
}

int rvvla_gen_random_rows_var(int64_t n, int kmin, int kmax, uint64_t seed, rvvla_gen_csr* out) {
  if (!out || n < 1 || kmin < 0 || kmax < kmin || kmax > n) return RVVLA_ERR_INVALID_ARG;
  std::vector<int64_t> len((size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    Rng rs(seed, 1, (uint64_t)r);
    len[(size_t)r] = kmin + rs.below(kmax - kmin + 1);
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < n; ++r) nnz += len[(size_t)r];
  if (nnz > 2147483647LL) return RVVLA_ERR_LENGTH;
  if (int rc = alloc_out(out, n, n, nnz)) return rc;
  out->rst[0] = 0;
  for (int64_t r = 0; r < n; ++r) out->rst[r + 1] = out->rst[r] + (int32_t)len[(size_t)r];
#pragma omp parallel
  {
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
      Rng rs(seed, 1, (uint64_t)r), rv(seed, 2, (uint64_t)r);
      rs.below(1);  // the length draw
      cols.clear();
      pick_distinct(cols, (size_t)len[(size_t)r], 0, n - 1, rs);
      const int64_t base = out->rst[r];
      for (int64_t j = 0; j < len[(size_t)r]; ++j) {
        out->col[base + j] = cols[(size_t)j];
        out->val[base + j] = rv.normal();
      }
    }
  }
  return RVVLA_OK;
}

int rvvla_gen_powerlaw(int64_t n, double alpha, int64_t maxlen, int64_t target_nnz, int64_t band,
                       uint64_t seed, rvvla_gen_csr* out) {
  if (!out || n < 1 || maxlen < 1 || maxlen > n || alpha <= 1.0 || band < 0 || target_nnz < n)
    return RVVLA_ERR_INVALID_ARG;
  // CDF of P(L = k) ~ k^-alpha on [1, maxlen]
  std::vector<double> cdf((size_t)maxlen);
  double acc = 0.0;
  for (int64_t k = 1; k <= maxlen; ++k) {
    acc += std::pow((double)k, -alpha);
    cdf[(size_t)(k - 1)] = acc;
  }
  for (auto& c : cdf) c /= acc;
  std::vector<int64_t> raw((size_t)n);
  double raw_sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : raw_sum)
  for (int64_t r = 0; r < n; ++r) {
    Rng rs(seed, 1, (uint64_t)r);
    const double u = rs.u01();
    raw[(size_t)r] = (int64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    if (raw[(size_t)r] > maxlen) raw[(size_t)r] = maxlen;
    raw_sum += (double)raw[(size_t)r];
  }
  // scale s such that sum_i clip(raw_i * s, 1, maxlen) == target_nnz
  // (bisection; the stochastic rounding below keeps that expectation)
  (void)raw_sum;
  double lo_s = 0.0, hi_s = (double)maxlen;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo_s + hi_s);
    double tot = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : tot)
    for (int64_t r = 0; r < n; ++r)
      tot += std::min((double)maxlen, std::max(1.0, (double)raw[(size_t)r] * mid));
    if (tot < (double)target_nnz) lo_s = mid; else hi_s = mid;
  }
  const double scale = 0.5 * (lo_s + hi_s);
  std::vector<int64_t> len((size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    Rng rs(seed, 1, (uint64_t)r);
    rs.u01();  // the length draw
    const double t = std::min((double)maxlen, std::max(1.0, (double)raw[(size_t)r] * scale));
    len[(size_t)r] = std::min<int64_t>(maxlen, (int64_t)std::floor(t + rs.u01()));
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < n; ++r) nnz += len[(size_t)r];
  if (nnz > 2147483647LL) return RVVLA_ERR_LENGTH;
  if (int rc = alloc_out(out, n, n, nnz)) return rc;
  out->rst[0] = 0;
  for (int64_t r = 0; r < n; ++r) out->rst[r + 1] = out->rst[r] + (int32_t)len[(size_t)r];
#pragma omp parallel
  {
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < n; ++r) {
      Rng rs(seed, 1, (uint64_t)r), rv(seed, 2, (uint64_t)r);
      rs.u01();
      rs.u01();  // length and rounding draws
      const int64_t L = len[(size_t)r];
      const int64_t lo = std::max<int64_t>(0, r - band), hi = std::min<int64_t>(n - 1, r + band);
      const int64_t near = std::min<int64_t>((L + 1) / 2, hi - lo + 1);
      cols.clear();
      pick_distinct(cols, (size_t)near, lo, hi, rs);
      top_up(cols, (size_t)L, 0, n - 1, rs);
      const int64_t base = out->rst[r];
      for (int64_t j = 0; j < L; ++j) {
        out->col[base + j] = cols[(size_t)j];
        out->val[base + j] = rv.normal();
      }
    }
  }
  return RVVLA_OK;
}

int rvvla_gen_uniform(int64_t count, double lo, double hi, uint64_t seed, uint64_t purpose,
                      double* v) {
  if (count < 0 || (count > 0 && !v)) return RVVLA_ERR_INVALID_ARG;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    Rng r(seed, purpose, (uint64_t)i);
    v[i] = lo + (hi - lo) * r.u01();
  }
  return RVVLA_OK;
}

}  // extern "C"
