// common.h -- shared parts of the accuracy and performance programs: matrix
// sources, format / dtype lists, device buffers, value narrowing.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "rvvla_hip.h"

namespace apps {

inline void ck(int rc, const std::string& what) {
  if (rc != RVVLA_OK) throw std::runtime_error(what + ": " + rvvla_strerror(rc));
}
inline void hk(hipError_t e, const std::string& what) {
  if (e != hipSuccess) throw std::runtime_error(what + ": " + hipGetErrorString(e));
}

// A matrix from a source spec: a .mtx or .bin file, "lap:G" (5-point
// Laplacian on G x G), "powerlaw:N[:NNZ]" (power-law rows, alpha 1.8) or
// "random:N:K" (diagonal + K-1 uniform columns).
struct Matrix {
  rvvla_gen_csr g{};
  std::string name;
  ~Matrix() { rvvla_gen_free(&g); }
};

inline bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

inline void load(const std::string& spec, Matrix& M) {
  M.name = spec;
  std::vector<std::string> parts;
  size_t p = 0;
  while (true) {
    const size_t q = spec.find(':', p);
    parts.push_back(spec.substr(p, q == std::string::npos ? std::string::npos : q - p));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  char code[4];
  if (ends_with(spec, ".mtx")) ck(rvvla_mtx_read(spec.c_str(), &M.g, code), "read " + spec);
  else if (ends_with(spec, ".bin")) ck(rvvla_bin_read(spec.c_str(), &M.g, code), "read " + spec);
  else if (parts[0] == "lap" && parts.size() == 2)
    ck(rvvla_gen_laplacian5(std::atoll(parts[1].c_str()), &M.g), spec);
  else if (parts[0] == "powerlaw" && parts.size() >= 2) {
    const long long n = std::atoll(parts[1].c_str());
    const long long nnz = parts.size() > 2 ? std::atoll(parts[2].c_str()) : 15 * n;
    ck(rvvla_gen_powerlaw(n, 1.8, 50000, nnz, 1024, 0, &M.g), spec);
  } else if (parts[0] == "random" && parts.size() == 3)
    ck(rvvla_gen_random_rows(std::atoll(parts[1].c_str()), std::atoi(parts[2].c_str()), 1, 0, 1.0,
                             0, &M.g),
       spec);
  else
    throw std::runtime_error("unknown matrix source " + spec +
                             " (a .mtx or .bin file, lap:G, powerlaw:N[:NNZ], random:N:K)");
}

inline std::vector<int> parse_formats(const std::string& s) {
  std::vector<int> f;
  if (s == "all") {
    for (int i = 0; i <= RVVLA_FMT_CVR; ++i) f.push_back(i);
    return f;
  }
  size_t p = 0;
  while (p <= s.size()) {
    const size_t q = s.find(',', p);
    const std::string t = s.substr(p, q == std::string::npos ? std::string::npos : q - p);
    int found = -1;
    for (int i = 0; i <= RVVLA_FMT_CVR; ++i)
      if (t == rvvla_format_name(i)) found = i;
    if (found < 0) throw std::runtime_error("unknown format " + t);
    f.push_back(found);
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return f;
}

inline std::vector<int> parse_dtypes(const std::string& s) {
  std::vector<int> d;
  size_t p = 0;
  while (p <= s.size()) {
    const size_t q = s.find(',', p);
    const std::string t = s.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (t == "f64") d.push_back(RVVLA_F64);
    else if (t == "f32") d.push_back(RVVLA_F32);
    else if (t == "bf16") d.push_back(RVVLA_BF16);
    else if (t == "f16") d.push_back(RVVLA_F16);
    else throw std::runtime_error("unknown dtype " + t + " (f64, f32, bf16, f16)");
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return d;
}

inline const char* dtype_name(int d) {
  return d == RVVLA_F64 ? "f64" : d == RVVLA_F32 ? "f32" : d == RVVLA_BF16 ? "bf16" : "f16";
}
inline size_t value_bytes(int d) { return d == RVVLA_F64 ? 8 : d == RVVLA_F32 ? 4 : 2; }
inline size_t vec_bytes(int d) { return d == RVVLA_F64 ? 8 : 4; }

// Host value array for a dtype (doubles for F64, floats otherwise).
struct Values {
  std::vector<double> d;
  std::vector<float> f;
  const void* ptr(int dtype) const { return dtype == RVVLA_F64 ? (const void*)d.data() : (const void*)f.data(); }
};
inline Values values_for(const rvvla_gen_csr& g, int dtype) {
  Values v;
  if (dtype == RVVLA_F64) v.d.assign(g.val, g.val + g.nnz);
  else v.f.assign(g.val, g.val + g.nnz);
  return v;
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { hk(hipMalloc(&p, bytes ? bytes : 8), "hipMalloc"); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Compulsory bytes of one SpMV with beta = 0.
inline double algo_bytes(const rvvla_gen_csr& g, int dtype) {
  return (double)g.nnz * (value_bytes(dtype) + 4) + (double)(g.m + 1) * 4 +
         (double)g.n * vec_bytes(dtype) + (double)g.m * vec_bytes(dtype);
}

}  // namespace apps
