// mtx_io.cpp -- Matrix Market and binary CSR files (host).
//
// The reference keeps an MM_typecode in spMtxCRS (reference
// include/CRS.h:8,17) and names a mtx-to-bin converter in its build options
// (reference cmake/TestEnvironment.cmake:62) but ships neither reader.  This
// file provides both, for the NIST Matrix Market exchange format
// (coordinate; real, integer or pattern; general, symmetric or
// skew-symmetric):
//  * the file is read whole and parsed in parallel: each thread takes a
//    byte range of the entry section starting at a line boundary, counts
//    its entry lines, then parses them into its slice of the COO arrays;
//  * symmetric and skew-symmetric files are expanded (the mirrored entry of
//    an off-diagonal (i, j, v) is (j, i, v) or (j, i, -v));
//  * rows are built by a counting sort, columns sorted within each row and
//    duplicate (i, j) entries summed, so the result is a canonical CSR.
// The binary file is a fixed header followed by the three CSR arrays, all
// little-endian:
//    0  char[8]  "RVVLACSR"
//    8  int32    version (1)
//   12  char[4]  MM_typecode
//   16  int64    m, n, nnz
//   40  int32    Rst[m + 1], int32 Col[nnz], float64 Val[nnz]
#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "internal.h"

namespace rvvla {
namespace {

struct File {
  std::FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

std::vector<char> read_all(const char* path) {
  File fh(path, "rb");
  if (!fh.f) throw Error(RVVLA_ERR_NOT_FOUND, std::string("cannot open ") + path);
  std::vector<char> buf;
  if (std::fseek(fh.f, 0, SEEK_END) == 0) {
    const long sz = std::ftell(fh.f);
    if (sz > 0) {
      buf.resize((size_t)sz);
      std::rewind(fh.f);
      if (std::fread(buf.data(), 1, buf.size(), fh.f) != buf.size())
        throw Error(RVVLA_ERR_INVALID_ARG, std::string("short read of ") + path);
    }
  }
  buf.push_back('\n');  // the last line always ends
  return buf;
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Parses an unsigned decimal at p; returns false when there is none.
bool parse_i64(const char*& p, const char* end, int64_t& v) {
  while (p < end && (*p == ' ' || *p == '\t')) ++p;
  if (p >= end || *p < '0' || *p > '9') return false;
  int64_t x = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    x = x * 10 + (*p - '0');
    if (x > (int64_t(1) << 53)) return false;
    ++p;
  }
  v = x;
  return true;
}

bool parse_f64(const char*& p, const char* end, double& v) {
  while (p < end && (*p == ' ' || *p == '\t')) ++p;
  if (p >= end || *p == '\n' || *p == '\r') return false;
  char* q = nullptr;
  v = std::strtod(p, &q);
  if (q == p) return false;
  p = q;
  return true;
}

const char* next_line(const char* p, const char* end) {
  const void* nl = std::memchr(p, '\n', (size_t)(end - p));
  return nl ? (const char*)nl + 1 : end;
}

bool blank_or_comment(const char* p, const char* end) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  return p >= end || *p == '\n' || *p == '%';
}

}  // namespace

// COO (0-based, any order, duplicates allowed) -> canonical CSR.
void coo_to_csr(int64_t m, int64_t n, std::vector<int64_t>& ri, std::vector<int64_t>& ci,
                std::vector<double>& v, rvvla_gen_csr* out) {
  const int64_t k = (int64_t)ri.size();
  std::vector<int64_t> cnt((size_t)m + 1, 0);
  for (int64_t e = 0; e < k; ++e) cnt[(size_t)ri[(size_t)e] + 1]++;
  for (int64_t r = 0; r < m; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  std::vector<int32_t> col((size_t)k);
  std::vector<double> val((size_t)k);
  for (int64_t e = 0; e < k; ++e) {
    const int64_t q = fill[(size_t)ri[(size_t)e]]++;
    col[(size_t)q] = (int32_t)ci[(size_t)e];
    val[(size_t)q] = v[(size_t)e];
  }
  ri.clear();
  ri.shrink_to_fit();
  ci.clear();
  ci.shrink_to_fit();
  v.clear();
  v.shrink_to_fit();
  // sort each row by column (stable: duplicates keep file order) and sum
  // duplicates; rows shrink in place, then compact
  std::vector<int64_t> keep((size_t)m, 0);
#pragma omp parallel
  {
    std::vector<int64_t> idx;
    std::vector<int32_t> c2;
    std::vector<double> v2;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < m; ++r) {
      const int64_t a = cnt[(size_t)r], b = cnt[(size_t)r + 1];
      if (b - a <= 1) {
        keep[(size_t)r] = b - a;
        continue;
      }
      idx.resize((size_t)(b - a));
      std::iota(idx.begin(), idx.end(), a);
      std::stable_sort(idx.begin(), idx.end(),
                       [&](int64_t x, int64_t y) { return col[(size_t)x] < col[(size_t)y]; });
      c2.clear();
      v2.clear();
      for (int64_t q : idx) {
        if (!c2.empty() && c2.back() == col[(size_t)q]) {
          v2.back() += val[(size_t)q];
        } else {
          c2.push_back(col[(size_t)q]);
          v2.push_back(val[(size_t)q]);
        }
      }
      std::copy(c2.begin(), c2.end(), col.begin() + a);
      std::copy(v2.begin(), v2.end(), val.begin() + a);
      keep[(size_t)r] = (int64_t)c2.size();
    }
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < m; ++r) nnz += keep[(size_t)r];
  if (nnz > 2147483647LL) throw Error(RVVLA_ERR_LENGTH, "matrix file: more than 2^31-1 entries");
  out->m = m;
  out->n = n;
  out->nnz = nnz;
  out->rst = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(m + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  out->val = (double*)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  if (!out->rst || !out->col || !out->val) {
    std::free(out->rst);
    std::free(out->col);
    std::free(out->val);
    out->rst = nullptr;
    out->col = nullptr;
    out->val = nullptr;
    throw Error(RVVLA_ERR_OUT_OF_MEMORY, "matrix file: out of memory");
  }
  int64_t w = 0;
  out->rst[0] = 0;
  for (int64_t r = 0; r < m; ++r) {
    const int64_t a = cnt[(size_t)r];
    for (int64_t q = 0; q < keep[(size_t)r]; ++q) {
      out->col[w] = col[(size_t)(a + q)];
      out->val[w] = val[(size_t)(a + q)];
      ++w;
    }
    out->rst[r + 1] = (int32_t)w;
  }
}

void mtx_read(const char* path, rvvla_gen_csr* out, char* matcode) {
  const std::vector<char> buf = read_all(path);
  const char* p = buf.data();
  const char* end = buf.data() + buf.size();
  // banner
  const char* l1 = next_line(p, end);
  std::string banner(p, l1);
  while (!banner.empty() && (banner.back() == '\n' || banner.back() == '\r')) banner.pop_back();
  char w[5][64] = {{0}};
  if (std::sscanf(banner.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5 ||
      lower(w[0]) != "%%matrixmarket" || lower(w[1]) != "matrix")
    throw Error(RVVLA_ERR_INVALID_ARG, "Matrix Market: bad banner line");
  const std::string fmt = lower(w[2]), field = lower(w[3]), sym = lower(w[4]);
  if (fmt != "coordinate")
    throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "Matrix Market: only the coordinate format is read");
  if (field != "real" && field != "integer" && field != "pattern" && field != "double")
    throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "Matrix Market: field " + field + " is not supported");
  if (sym != "general" && sym != "symmetric" && sym != "skew-symmetric")
    throw Error(RVVLA_ERR_UNSUPPORTED_SPMV, "Matrix Market: symmetry " + sym + " is not supported");
  const bool pattern = field == "pattern";
  const int mirror = sym == "general" ? 0 : sym == "symmetric" ? 1 : -1;
  if (matcode) {
    matcode[0] = 'M';
    matcode[1] = 'C';
    matcode[2] = pattern ? 'P' : field == "integer" ? 'I' : 'R';
    matcode[3] = mirror == 0 ? 'G' : mirror > 0 ? 'S' : 'K';
  }
  // size line
  p = l1;
  while (p < end && blank_or_comment(p, end)) p = next_line(p, end);
  int64_t m = 0, n = 0, k = 0;
  {
    const char* q = p;
    if (!parse_i64(q, end, m) || !parse_i64(q, end, n) || !parse_i64(q, end, k))
      throw Error(RVVLA_ERR_INVALID_ARG, "Matrix Market: bad size line");
    p = next_line(p, end);
  }
  if (m > 2147483646LL || n > 2147483647LL)
    throw Error(RVVLA_ERR_LENGTH, "Matrix Market: dimensions beyond the int32 range");
  if (mirror && m != n)
    throw Error(RVVLA_ERR_INVALID_ARG, "Matrix Market: a symmetric matrix must be square");
  // entry section split among threads at line starts
  const int T = std::max(1, std::min(omp_get_max_threads(), (int)((end - p) / (1 << 16)) + 1));
  std::vector<const char*> cut((size_t)T + 1);
  cut[0] = p;
  cut[(size_t)T] = end;
  for (int t = 1; t < T; ++t) {
    const char* c = p + (end - p) * t / T;
    if (c > p && c[-1] != '\n') c = next_line(c, end);
    cut[(size_t)t] = std::max(c, cut[(size_t)t - 1]);
  }
  std::vector<int64_t> lines((size_t)T + 1, 0);
#pragma omp parallel for num_threads(T)
  for (int t = 0; t < T; ++t) {
    int64_t c = 0;
    for (const char* q = cut[(size_t)t]; q < cut[(size_t)t + 1]; q = next_line(q, cut[(size_t)t + 1]))
      c += !blank_or_comment(q, cut[(size_t)t + 1]);
    lines[(size_t)t + 1] = c;
  }
  for (int t = 0; t < T; ++t) lines[(size_t)t + 1] += lines[(size_t)t];
  if (lines[(size_t)T] != k)
    throw Error(RVVLA_ERR_INVALID_ARG, "Matrix Market: " + std::to_string(lines[(size_t)T]) +
                                           " entry lines, the size line says " + std::to_string(k));
  std::vector<int64_t> ri((size_t)k), ci((size_t)k);
  std::vector<double> v((size_t)k);
  int bad = 0;
#pragma omp parallel for num_threads(T) reduction(| : bad)
  for (int t = 0; t < T; ++t) {
    int64_t e = lines[(size_t)t];
    const char* lim = cut[(size_t)t + 1];
    for (const char* q = cut[(size_t)t]; q < lim; q = next_line(q, lim)) {
      if (blank_or_comment(q, lim)) continue;
      const char* s = q;
      int64_t i = 0, j = 0;
      double x = 1.0;
      if (!parse_i64(s, lim, i) || !parse_i64(s, lim, j) || (!pattern && !parse_f64(s, lim, x)) ||
          i < 1 || i > m || j < 1 || j > n) {
        bad = 1;
        break;
      }
      ri[(size_t)e] = i - 1;
      ci[(size_t)e] = j - 1;
      v[(size_t)e] = x;
      ++e;
    }
  }
  if (bad) throw Error(RVVLA_ERR_INVALID_ARG, "Matrix Market: malformed or out-of-range entry");
  if (mirror) {
    int64_t off = 0;
    for (int64_t e = 0; e < k; ++e) off += ri[(size_t)e] != ci[(size_t)e];
    ri.reserve((size_t)(k + off));
    ci.reserve((size_t)(k + off));
    v.reserve((size_t)(k + off));
    for (int64_t e = 0; e < k; ++e)
      if (ri[(size_t)e] != ci[(size_t)e]) {
        ri.push_back(ci[(size_t)e]);
        ci.push_back(ri[(size_t)e]);
        v.push_back(mirror > 0 ? v[(size_t)e] : -v[(size_t)e]);
      }
  }
  coo_to_csr(m, n, ri, ci, v, out);
}

void bin_write(const char* path, const rvvla_gen_csr* a, const char* matcode) {
  File fh(path, "wb");
  if (!fh.f) throw Error(RVVLA_ERR_NOT_FOUND, std::string("cannot create ") + path);
  const int32_t version = 1;
  char code[4] = {'M', 'C', 'R', 'G'};
  if (matcode) std::memcpy(code, matcode, 4);
  const int64_t dims[3] = {a->m, a->n, a->nnz};
  bool ok = std::fwrite("RVVLACSR", 1, 8, fh.f) == 8 && std::fwrite(&version, 4, 1, fh.f) == 1 &&
            std::fwrite(code, 1, 4, fh.f) == 4 && std::fwrite(dims, 8, 3, fh.f) == 3 &&
            std::fwrite(a->rst, 4, (size_t)a->m + 1, fh.f) == (size_t)a->m + 1;
  if (ok && a->nnz > 0)
    ok = std::fwrite(a->col, 4, (size_t)a->nnz, fh.f) == (size_t)a->nnz &&
         std::fwrite(a->val, 8, (size_t)a->nnz, fh.f) == (size_t)a->nnz;
  if (!ok) throw Error(RVVLA_ERR_INVALID_ARG, std::string("short write to ") + path);
}

void bin_read(const char* path, rvvla_gen_csr* out, char* matcode) {
  File fh(path, "rb");
  if (!fh.f) throw Error(RVVLA_ERR_NOT_FOUND, std::string("cannot open ") + path);
  char magic[8];
  int32_t version = 0;
  char code[4];
  int64_t dims[3];
  if (std::fread(magic, 1, 8, fh.f) != 8 || std::memcmp(magic, "RVVLACSR", 8) != 0 ||
      std::fread(&version, 4, 1, fh.f) != 1 || version != 1 ||
      std::fread(code, 1, 4, fh.f) != 4 || std::fread(dims, 8, 3, fh.f) != 3)
    throw Error(RVVLA_ERR_INVALID_ARG, std::string("not a binary CSR file: ") + path);
  const int64_t m = dims[0], n = dims[1], nnz = dims[2];
  if (m < 0 || n < 0 || nnz < 0 || m > 2147483646LL || n > 2147483647LL || nnz > 2147483647LL)
    throw Error(RVVLA_ERR_INVALID_ARG, "binary CSR file: bad dimensions");
  out->m = m;
  out->n = n;
  out->nnz = nnz;
  out->rst = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(m + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  out->val = (double*)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  bool ok = out->rst && out->col && out->val &&
            std::fread(out->rst, 4, (size_t)m + 1, fh.f) == (size_t)m + 1;
  if (ok && nnz > 0)
    ok = std::fread(out->col, 4, (size_t)nnz, fh.f) == (size_t)nnz &&
         std::fread(out->val, 8, (size_t)nnz, fh.f) == (size_t)nnz;
  if (ok) {
    try {
      check_csr_public(m, n, nnz, out->rst, out->col);
    } catch (const Error&) {
      ok = false;
    }
  }
  if (!ok) {
    std::free(out->rst);
    std::free(out->col);
    std::free(out->val);
    out->rst = nullptr;
    out->col = nullptr;
    out->val = nullptr;
    throw Error(RVVLA_ERR_INVALID_ARG, std::string("truncated or invalid binary CSR file: ") + path);
  }
  if (matcode) std::memcpy(matcode, code, 4);
}

}  // namespace rvvla

extern "C" {

static int io_guard(void (*body)(void*), void* ctx) {
  try {
    body(ctx);
    return RVVLA_OK;
  } catch (const rvvla::Error& e) {
    rvvla::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    rvvla::set_last_error("out of host memory");
    return RVVLA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    rvvla::set_last_error(e.what());
    return RVVLA_ERR_LOGIC;
  }
}

struct IoCtx {
  const char* path;
  rvvla_gen_csr* out;
  const rvvla_gen_csr* in;
  char* code;
  const char* ccode;
};

int rvvla_mtx_read(const char* path, rvvla_gen_csr* out, char matcode[4]) {
  if (!path || !out) return RVVLA_ERR_INVALID_ARG;
  out->rst = nullptr;
  out->col = nullptr;
  out->val = nullptr;
  IoCtx c{path, out, nullptr, matcode, nullptr};
  return io_guard([](void* p) {
    IoCtx* c = (IoCtx*)p;
    rvvla::mtx_read(c->path, c->out, c->code);
  }, &c);
}

int rvvla_bin_write(const char* path, const rvvla_gen_csr* a, const char matcode[4]) {
  if (!path || !a || !a->rst || (a->nnz > 0 && (!a->col || !a->val))) return RVVLA_ERR_INVALID_ARG;
  IoCtx c{path, nullptr, a, nullptr, matcode};
  return io_guard([](void* p) {
    IoCtx* c = (IoCtx*)p;
    rvvla::check_csr_public(c->in->m, c->in->n, c->in->nnz, c->in->rst, c->in->col);
    rvvla::bin_write(c->path, c->in, c->ccode);
  }, &c);
}

int rvvla_bin_read(const char* path, rvvla_gen_csr* out, char matcode[4]) {
  if (!path || !out) return RVVLA_ERR_INVALID_ARG;
  out->rst = nullptr;
  out->col = nullptr;
  out->val = nullptr;
  IoCtx c{path, out, nullptr, matcode, nullptr};
  return io_guard([](void* p) {
    IoCtx* c = (IoCtx*)p;
    rvvla::bin_read(c->path, c->out, c->code);
  }, &c);
}

}  // extern "C"
