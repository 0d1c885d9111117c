// oracle.cpp -- CPU restatement of the reference SpMV path.  TEST
// INFRASTRUCTURE ONLY: loaded by tests/, by __graft_entry__.smoke() and by
// bench.py's CPU-baseline leg, never by the library.
//
// Every function names the reference lines it restates.  It is compiled the
// way the reference scalar library is (g++ -O3 -fopenmp -fno-tree-vectorize,
// reference src/scalar/CMakeLists.txt:17-18), so its sums round exactly as
// the reference's scalar kernels do on x86-64.
//
// Pinning (what checks each function against something independent):
//  * orc_csr_spmv_*      : bit-equal to the compiled reference scalar CSR
//                          (oracle/_ref) and to golden hex vectors; Laplacian
//                          closed forms; dense brute force on tiny inputs.
//  * orc_hcsr_convert_*  : bit-equal to the reference converter crs_to_HCSR
//                          (oracle/_ref), to golden fixtures and to the
//                          worked 6x6 example.
//  * orc_hcsr_spmv_*     : bit-equal to the reference scalar HCSR kernel.
//  * orc_sell_build_*    : no reference builder exists; its layout is fed to
//                          the reference scalar SELL kernel (oracle/_ref) and
//                          must reproduce the CSR result; sortedness and
//                          permutation invariants.
//  * orc_sell_spmv_*     : bit-equal to the reference scalar SELL kernel.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

namespace {

// A converted structure: named arrays (raw bytes + element size) and scalars.
struct Obj {
  std::map<std::string, std::pair<std::vector<char>, int>> arrays;
  std::map<std::string, double> scalars;
  template <class T>
  void put(const std::string& k, const std::vector<T>& v) {
    std::vector<char> b(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(b.data(), v.data(), b.size());
    arrays[k] = {std::move(b), (int)sizeof(T)};
  }
  template <class T>
  const T* get(const std::string& k) const {
    return (const T*)arrays.at(k).first.data();
  }
  int64_t count(const std::string& k) const {
    const auto& a = arrays.at(k);
    return (int64_t)a.first.size() / a.second;
  }
};

// ---------------------------------------------------------------- CSR
// reference src/scalar/spmv_mtx_crs.cpp:22-46 (double): tmp summed in
// entry order, y = tmp * alpha + beta * y, rows in dynamic chunks of 256.
void csr_f64(int64_t m, const int* rst, const int* col, const double* val, const double* b,
             double alpha, double beta, double* y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < m; i++) {
    double tmp = 0.0;
    for (int j = rst[i]; j < rst[i + 1]; j++) tmp += val[j] * b[col[j]];
    y[i] = tmp * alpha + beta * y[i];
  }
}

// reference src/scalar/spmv_mtx_crs.cpp:48-71 (float): y = alpha * tmp + beta * y.
void csr_f32(int64_t m, const int* rst, const int* col, const float* val, const float* b,
             float alpha, float beta, float* y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < m; i++) {
    float tmp = 0.0f;
    for (int j = rst[i]; j < rst[i + 1]; j++) tmp += val[j] * b[col[j]];
    y[i] = alpha * tmp + beta * y[i];
  }
}

// ---------------------------------------------------------------- HCSR
// reference src/scalar/crs_to_hcsr.cpp:30-135, written serially.
template <class T>
Obj* hcsr_convert(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col, const T* val,
                  int64_t R, int64_t C, double p, int* err) {
  *err = 0;
  if (R <= 0 || C <= 0) { *err = -10; return nullptr; }
  const int64_t bm = (m + R - 1) / R, bn = (n + C - 1) / C, nb = bm * bn;
  std::vector<int> block_ptr((size_t)nb + 1, 0);
  std::vector<uint32_t> row_offset((size_t)nb + 1, 0);
  std::vector<int> kind((size_t)nb, 0);
  for (int64_t br = 0; br < bm; ++br) {
    const int64_t r0 = br * R, r1 = std::min(m, r0 + R);
    std::vector<int64_t> cnt((size_t)bn, 0);
    for (int64_t i = r0; i < r1; ++i)
      for (int k = rst[i]; k < rst[i + 1]; ++k) cnt[(size_t)(col[k] / C)]++;
    for (int64_t j = 0; j < bn; ++j) {
      const int64_t b = br * bn + j;
      block_ptr[(size_t)b + 1] = (int)cnt[(size_t)j];
      if (cnt[(size_t)j] == 0) kind[(size_t)b] = 0;
      else if (double(r1 - r0) * p < double(cnt[(size_t)j])) {
        kind[(size_t)b] = 1;
        row_offset[(size_t)b + 1] = (uint32_t)(r1 - r0 + 1);
      } else {
        kind[(size_t)b] = 2;
        row_offset[(size_t)b + 1] = (uint32_t)cnt[(size_t)j];
      }
    }
  }
  for (int64_t b = 1; b <= nb; ++b) {
    block_ptr[(size_t)b] += block_ptr[(size_t)b - 1];
    row_offset[(size_t)b] += row_offset[(size_t)b - 1];
  }
  const uint64_t total_rp = row_offset[(size_t)nb];
  if (total_rp > (1u << 30) - 1) { *err = -11; return nullptr; }
  if (block_ptr[(size_t)nb] != nnz) { *err = -12; return nullptr; }
  std::vector<uint32_t> row_ptr((size_t)total_rp, 0);
  std::vector<uint16_t> col_idx((size_t)nnz);
  std::vector<T> values((size_t)nnz);
  for (int64_t br = 0; br < bm; ++br) {
    const int64_t r0 = br * R, r1 = std::min(m, r0 + R);
    std::vector<int64_t> pos((size_t)bn);
    for (int64_t j = 0; j < bn; ++j) pos[(size_t)j] = block_ptr[(size_t)(br * bn + j)];
    for (int64_t i = r0; i < r1; ++i) {
      for (int k = rst[i]; k < rst[i + 1]; ++k) {
        const int64_t j = col[k] / C, b = br * bn + j, q = pos[(size_t)j]++;
        col_idx[(size_t)q] = (uint16_t)(col[k] - j * C);
        values[(size_t)q] = val[k];
        if (kind[(size_t)b] == 1) row_ptr[row_offset[(size_t)b] + (size_t)(i - r0) + 1]++;
        else row_ptr[row_offset[(size_t)b] + (size_t)(q - block_ptr[(size_t)b])] = (uint32_t)(i - r0);
      }
    }
    for (int64_t j = 0; j < bn; ++j) {
      const int64_t b = br * bn + j;
      if (kind[(size_t)b] != 1) continue;
      for (int64_t li = 0; li < r1 - r0; ++li)
        row_ptr[row_offset[(size_t)b] + (size_t)li + 1] += row_ptr[row_offset[(size_t)b] + (size_t)li];
    }
  }
  for (int64_t b = 0; b < nb; ++b) row_offset[(size_t)b] |= (uint32_t)kind[(size_t)b] << 30;
  Obj* o = new Obj;
  o->put("block_ptr", block_ptr);
  o->put("row_offset", row_offset);
  o->put("row_ptr", row_ptr);
  o->put("col_idx", col_idx);
  o->put("values", values);
  o->scalars = {{"R", (double)R}, {"C", (double)C}, {"m", (double)m}, {"n", (double)n},
                {"nz", (double)nnz}, {"b_m", (double)bm}, {"b_n", (double)bn}};
  return o;
}

// reference src/scalar/spmv_mtx_hcsr.cpp:22-102: per block-row, y *= beta,
// then every block column in order (CSR kind: y[i] += alpha * row dot;
// any other kind: COO scatter y[row] += alpha * v * b, also for empty blocks
// where it adds nothing).
template <class T>
void hcsr_spmv(const Obj* o, const T* b, T alpha, T beta, T* y) {
  const int64_t m = (int64_t)o->scalars.at("m"), n = (int64_t)o->scalars.at("n");
  const int64_t R = (int64_t)o->scalars.at("R"), C = (int64_t)o->scalars.at("C");
  const int64_t bn = (int64_t)o->scalars.at("b_n");
  const int* bp = o->get<int>("block_ptr");
  const uint32_t* ro = o->get<uint32_t>("row_offset");
  const uint32_t* rp = o->get<uint32_t>("row_ptr");
  const uint16_t* ci = o->get<uint16_t>("col_idx");
  const T* v = o->get<T>("values");
#pragma omp parallel for schedule(dynamic)
  for (int64_t i = 0; i < m; i += R) {
    const int64_t maxr = std::min(R, m - i);
    for (int64_t r = 0; r < maxr; ++r) y[i + r] *= beta;
    for (int64_t j = 0; j < n; j += C) {
      const int64_t blk = j / C + (i / R) * bn;
      const int start = bp[blk];
      const uint32_t key = ro[blk] >> 30, off = ro[blk] & 0x3FFFFFFFu;
      const int nz = bp[blk + 1] - start;
      const uint32_t* lrp = rp + off;
      const uint16_t* lc = ci + start;
      const T* lv = v + start;
      if (key == 1) {
        for (int64_t r = 0; r < maxr; ++r) {
          T tmp = 0;
          for (uint32_t k = lrp[r]; k < lrp[r + 1]; ++k) tmp += lv[k] * b[j + lc[k]];
          y[i + r] += alpha * tmp;
        }
      } else {
        for (int k = 0; k < nz; ++k) y[i + lrp[k]] += alpha * lv[k] * b[j + lc[k]];
      }
    }
  }
}

// ---------------------------------------------------------------- SELL-C-sigma
// Builder (no reference builder exists): stable descending sort by row
// length inside windows of sigma rows (Kreutzer et al. 2014); slices of C
// sorted rows; slot (j, k) at Cs + j*Bs + k; padding = value 0 with the
// row's last column (0 for an empty row).
template <class T>
Obj* sell_build(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col, const T* val,
                int C, int sigma) {
  std::vector<uint32_t> perm((size_t)m);
  for (int64_t i = 0; i < m; ++i) perm[(size_t)i] = (uint32_t)i;
  for (int64_t w = 0; w < m; w += sigma) {
    const int64_t e = std::min<int64_t>(m, w + sigma);
    std::stable_sort(perm.begin() + w, perm.begin() + e, [&](uint32_t a, uint32_t b) {
      return (rst[a + 1] - rst[a]) > (rst[b + 1] - rst[b]);
    });
  }
  const int64_t nb = (m + C - 1) / C;
  std::vector<int> Cl((size_t)nb), Cs((size_t)nb), Bs((size_t)nb), SBs((size_t)nb);
  int64_t slots = 0;
  for (int64_t i = 0; i < nb; ++i) {
    SBs[(size_t)i] = (int)(i * C);
    Bs[(size_t)i] = (int)std::min<int64_t>(C, m - i * C);
    int w = 0;
    for (int k = 0; k < Bs[(size_t)i]; ++k) {
      const uint32_t r = perm[(size_t)(i * C + k)];
      w = std::max(w, rst[r + 1] - rst[r]);
    }
    Cl[(size_t)i] = w;
    Cs[(size_t)i] = (int)slots;
    slots += (int64_t)w * Bs[(size_t)i];
  }
  std::vector<uint32_t> vcol((size_t)slots, 0);
  std::vector<T> vval((size_t)slots, T(0));
  for (int64_t i = 0; i < nb; ++i)
    for (int k = 0; k < Bs[(size_t)i]; ++k) {
      const uint32_t r = perm[(size_t)(SBs[(size_t)i] + k)];
      const int len = rst[r + 1] - rst[r];
      for (int j = 0; j < Cl[(size_t)i]; ++j) {
        const size_t s = (size_t)Cs[(size_t)i] + (size_t)j * Bs[(size_t)i] + k;
        if (j < len) {
          vcol[s] = (uint32_t)col[rst[r] + j];
          vval[s] = val[rst[r] + j];
        } else {
          vcol[s] = len > 0 ? (uint32_t)col[rst[r] + len - 1] : 0u;
        }
      }
    }
  Obj* o = new Obj;
  o->put("vCl", Cl);
  o->put("vCs", Cs);
  o->put("vBs", Bs);
  o->put("vSBs", SBs);
  o->put("vCol", vcol);
  o->put("vPerm", perm);
  o->put("vVal", vval);
  o->scalars = {{"m", (double)m}, {"n", (double)n}, {"nz", (double)nnz}, {"cnt_b", (double)nb},
                {"C", (double)C}, {"Sigma", (double)sigma}};
  return o;
}

// reference src/scalar/spmv_mtx_sell_c_sigma.cpp:22-149 (all stages):
// y_s[i] = y[perm[i]]; per slice y_tmp[k] += Val*b[Col] over slots j;
// y_s = y_tmp*alpha + beta*y_s; y[perm[i]] = y_s[i].  Takes the arrays
// directly so any SELL layout (the library's or this oracle's) can be fed.
template <class T>
void sell_spmv(int64_t m, int64_t nb, const int* Cl, const int* Cs, const int* Bs, const int* SBs,
               const uint32_t* Col, const uint32_t* Perm, const T* Val, const T* b, T alpha,
               T beta, T* y) {
  std::vector<T> ys((size_t)m);
  for (int64_t i = 0; i < m; ++i) ys[(size_t)i] = y[Perm[i]];
  std::vector<T> acc(64 > (nb ? Bs[0] : 0) ? 64 : Bs[0]);
  for (int64_t i = 0; i < nb; ++i) {
    if ((int)acc.size() < Bs[i]) acc.resize((size_t)Bs[i]);
    for (int k = 0; k < Bs[i]; ++k) acc[(size_t)k] = 0;
    for (int j = 0; j < Cl[i]; ++j)
      for (int k = 0; k < Bs[i]; ++k) {
        const int64_t s = (int64_t)Cs[i] + (int64_t)j * Bs[i] + k;
        acc[(size_t)k] += Val[s] * b[Col[s]];
      }
    for (int k = 0; k < Bs[i]; ++k) {
      T& yy = ys[(size_t)(SBs[i] + k)];
      yy = acc[(size_t)k] * alpha + beta * yy;
    }
  }
  for (int64_t i = 0; i < m; ++i) y[Perm[i]] = ys[(size_t)i];
}

// ---------------------------------------------------------------- CSR5
// Builder (the reference declares spMtxCSR5::inputCSR/asCSR5 but ships no
// definition, reference include/CSR5.h:61-65): Liu & Vinter's tile layout as
// the reference vector kernel reads it (src/rvv/spmv_mtx_csr5.cpp:145-294).
// Written element by element from a row-of-entry table.
template <class T>
Obj* csr5_build(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col, const T* val,
                int omega, int sigma) {
  const int64_t W = (int64_t)omega * sigma;
  const int64_t p = nnz > 0 ? (nnz + W - 1) / W : 1;
  auto clog2 = [](int64_t v) { int b = 0; while ((int64_t(1) << b) < v) ++b; return b; };
  const int by = clog2(W), bs = clog2(omega), dl = (by + bs + sigma + 31) / 32;
  const int f0 = 32 - by - bs;
  std::vector<int64_t> row_of((size_t)nnz);
  for (int64_t r = 0; r < m; ++r)
    for (int e = rst[r]; e < rst[r + 1]; ++e) row_of[(size_t)e] = r;
  auto starts = [&](int64_t e) { return rst[row_of[(size_t)e]] == e; };
  std::vector<uint32_t> tp((size_t)p + 1);
  for (int64_t t = 0; t <= p; ++t) tp[(size_t)t] = t * W < nnz ? (uint32_t)row_of[(size_t)(t * W)] : (uint32_t)m;
  const int64_t nt = p - 1;
  std::vector<uint32_t> desc((size_t)std::max<int64_t>(nt, 0) * omega * dl, 0);
  std::vector<uint32_t> eptr((size_t)p + 1, 0), eoff;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t r0 = tp[(size_t)t], r1 = tp[(size_t)t + 1];
    bool has_empty = false;
    for (int64_t r = r0 + 1; r < r1; ++r) has_empty |= rst[r] == rst[r + 1];
    if (has_empty) tp[(size_t)t] |= 0x80000000u;
    eptr[(size_t)t] = (uint32_t)eoff.size();
    // flags and counts
    std::vector<int> cntl((size_t)omega, 0);
    for (int l = 0; l < omega; ++l)
      for (int i = 0; i < sigma; ++i) {
        const int64_t e = t * W + (int64_t)l * sigma + i;
        if (!starts(e)) continue;
        const int64_t w = i < f0 ? 0 : 1 + (i - f0) / 32;
        const int bit = i < f0 ? 31 - (32 - f0) - i : 31 - (i - f0) % 32;
        desc[(size_t)(t * omega * dl + w * omega + l)] |= 1u << bit;
        if (l == 0 && i == 0) continue;
        cntl[(size_t)l]++;
        if (has_empty) eoff.push_back((uint32_t)(row_of[(size_t)e] - r0 - 1));
      }
    uint32_t before = 0;
    for (int l = 0; l < omega; ++l) {
      int nx = l + 1;
      while (nx < omega && cntl[(size_t)nx] == 0) ++nx;
      const uint32_t seg = (uint32_t)(nx - l - 1);
      desc[(size_t)(t * omega * dl + l)] |= (before << (32 - by)) | (seg << (32 - by - bs));
      before += (uint32_t)cntl[(size_t)l];
    }
  }
  if (nt > 0) eptr[(size_t)nt] = (uint32_t)eoff.size();
  std::vector<uint32_t> c2((size_t)nnz);
  std::vector<T> v2((size_t)nnz);
  for (int64_t e = 0; e < nnz; ++e) {
    int64_t d = e;
    if (e < nt * W) {
      const int64_t t = e / W, l = (e % W) / sigma, i = (e % W) % sigma;
      d = t * W + i * omega + l;
    }
    c2[(size_t)d] = (uint32_t)col[e];
    v2[(size_t)d] = val[e];
  }
  Obj* o = new Obj;
  o->put("_csr_row_pointer", std::vector<int>(rst, rst + m + 1));
  o->put("_csr_column_index", c2);
  o->put("_csr_value", v2);
  o->put("_csr5_tile_pointer", tp);
  o->put("_csr5_tile_descriptor", desc);
  o->put("_csr5_tile_empty_offset_pointer", eptr);
  o->put("_csr5_tile_empty_offset", eoff);
  o->scalars = {{"_m", (double)m}, {"_n", (double)n}, {"_nnz", (double)nnz}, {"_p", (double)p},
                {"_omega", (double)omega}, {"_sigma", (double)sigma}, {"_bit_y_offset", (double)by},
                {"_bit_seg_offset", (double)bs}, {"_tile_desc_len", (double)dl},
                {"_tail_tile_start", (double)(tp[(size_t)p - 1] & 0x7FFFFFFFu)},
                {"_num_offsets", (double)eoff.size()}};
  return o;
}

// The reference vector CSR5 multiply restated lane by lane for one thread
// (reference src/rvv/spmv_mtx_csr5.cpp:498-517 entry: y *= beta;
// :85-296 compute; :318-335 calibrate; :337-365 tail).  Lane vectors are
// arrays of omega; v_incolumn starts false (the reference reads it before
// its first assignment); the simple-loop tile sums its omega*sigma entries
// in storage order.
template <class T>
void csr5_spmv_refsem(int64_t m, int64_t p, int omega, int sigma, int by, int bs, int dl,
                      int tail_start, const int* rst, const uint32_t* col, const T* val,
                      const uint32_t* tp, const uint32_t* desc, const uint32_t* eptr,
                      const uint32_t* eoff, const T* x, T alpha, T beta, T* y) {
  for (int64_t i = 0; i < m; ++i) y[i] *= beta;
  const int64_t W = (int64_t)omega * sigma;
  T calibrator = 0;
  const uint32_t thread_first_row = p - 1 > 0 ? (tp[0] & 0x7FFFFFFFu) : 0;
  std::vector<char> incolumn((size_t)omega, 0), local((size_t)omega), red((size_t)omega);
  std::vector<uint32_t> yoff((size_t)omega), segoff((size_t)omega), dsc((size_t)omega),
      start((size_t)omega), stop((size_t)omega);
  std::vector<T> sum((size_t)omega), first((size_t)omega), last((size_t)omega), tmp((size_t)omega),
      scan((size_t)omega);
  auto yidx = [&](bool empty, uint32_t off_ptr, uint32_t yo) { return empty ? eoff[off_ptr + yo] : yo; };
  for (int64_t t = 0; t < p - 1; ++t) {
    const int64_t base = t * W;
    uint32_t first_row = tp[t];
    const uint32_t last_row = tp[t + 1] & 0x7FFFFFFFu;
    if (first_row == last_row) {
      const bool row_starts_here = (desc[(size_t)(t * omega * dl)] >> (31 - (by + bs))) & 1u;
      T s = 0;
      for (int64_t k = 0; k < W; ++k) s += val[base + k] * x[col[base + k]];
      s *= alpha;
      if (first_row == thread_first_row && !row_starts_here) calibrator += s;
      else y[first_row] += s;
      continue;
    }
    const bool empty = (first_row >> 31) & 1u;
    first_row &= 0x7FFFFFFFu;
    T* yl = y + first_row + 1;
    const uint32_t off_ptr = empty ? eptr[t] : 0;
    const uint32_t* d = desc + t * omega * dl;
    for (int l = 0; l < omega; ++l) {
      first[(size_t)l] = 0;
      stop[(size_t)l] = 0;
      dsc[(size_t)l] = d[l];
      yoff[(size_t)l] = dsc[(size_t)l] >> (32 - by);
      segoff[(size_t)l] = (dsc[(size_t)l] << by) >> (32 - bs);
      dsc[(size_t)l] <<= (by + bs);
      local[(size_t)l] = (dsc[(size_t)l] >> 31) == 1u;
    }
    const bool first_incolumn = incolumn[0] != 0;
    const bool first_all_incolumn = (t == 0) ? first_incolumn : false;
    local[0] = 1;
    for (int l = 0; l < omega; ++l) {
      start[(size_t)l] = local[(size_t)l] ? 0u : 1u;
      incolumn[(size_t)l] = local[(size_t)l] && l != 0;
      sum[(size_t)l] = val[base + l] * x[col[base + l]];
    }
    int ly = 0;
    for (int i = 1; i < sigma; ++i) {
      int norm_i = i - (32 - by - bs);
      if (norm_i % 32 == 0) {
        ++ly;
        for (int l = 0; l < omega; ++l) dsc[(size_t)l] = d[ly * omega + l];
      }
      norm_i = (ly == 0) ? i : norm_i % 32;
      const int off = 31 - norm_i;
      bool any = false;
      for (int l = 0; l < omega; ++l) {
        local[(size_t)l] = (dsc[(size_t)l] >> off) & 1u;
        any |= local[(size_t)l] != 0;
      }
      if (any) {
        for (int l = 0; l < omega; ++l) {
          const bool green = incolumn[(size_t)l] && local[(size_t)l];
          if (green) {
            yl[yidx(empty, off_ptr, yoff[(size_t)l])] += alpha * sum[(size_t)l];
            yoff[(size_t)l] += 1;
          }
          red[(size_t)l] = local[(size_t)l] && !incolumn[(size_t)l];
          if (red[(size_t)l]) first[(size_t)l] = sum[(size_t)l];
          if (local[(size_t)l]) sum[(size_t)l] = 0;
          incolumn[(size_t)l] |= local[(size_t)l];
          stop[(size_t)l] += local[(size_t)l] ? 1u : 0u;
        }
      }
      for (int l = 0; l < omega; ++l) sum[(size_t)l] += val[base + (int64_t)i * omega + l] * x[col[base + (int64_t)i * omega + l]];
    }
    for (int l = 0; l < omega; ++l) {
      first[(size_t)l] = incolumn[(size_t)l] ? first[(size_t)l] : sum[(size_t)l];
      last[(size_t)l] = sum[(size_t)l];
    }
    for (int l = 0; l < omega; ++l) tmp[(size_t)l] = start[(size_t)l] == 0 ? T(0) : first[(size_t)l];
    for (int l = 0; l < omega; ++l) scan[(size_t)l] = l + 1 < omega ? tmp[(size_t)l + 1] : T(0);
    const std::vector<T> shifted = scan;
    for (int off = 1; off < omega; off <<= 1) {
      std::vector<T> prev = scan;
      for (int l = off; l < omega; ++l) scan[(size_t)l] = prev[(size_t)l] + prev[(size_t)(l - off)];
    }
    for (int l = 0; l < omega; ++l) {
      const uint32_t g = segoff[(size_t)l] + (uint32_t)l;
      const T s = (scan[g] - scan[(size_t)l]) + shifted[(size_t)l];
      if (start[(size_t)l] <= stop[(size_t)l]) last[(size_t)l] = last[(size_t)l] + s;
    }
    for (int l = 0; l < omega; ++l)
      if (incolumn[(size_t)l]) yl[yidx(empty, off_ptr, yoff[(size_t)l])] += alpha * last[(size_t)l];
    T add = incolumn[0] ? first[0] : last[0];
    add *= alpha;
    if (first_row == thread_first_row && !first_all_incolumn) calibrator += add;
    else y[first_row] += add;
  }
  if (p - 1 > 0) y[tp[0] & 0x7FFFFFFFu] += calibrator;
  const int64_t tail0 = (p - 1) * W;
  for (int64_t r = tail_start; r < m; ++r) {
    const int64_t e0 = r == tail_start ? tail0 : rst[r];
    T s = 0;
    for (int64_t e = e0; e < rst[r + 1]; ++e) s += val[e] * x[col[e]];
    s *= alpha;
    y[r] += s;
  }
}

// The reference vector VHCC multiply restated lane by lane, for any vector
// width (reference src/rvv/spmv_vhcc.cpp:73-162 kernel, :340-394 MULT stage:
// tmp_y of pad_rows zeroed, partitions in order, then
// result[last_row] += result[overflow_row] per partition, then
// y = y*beta + alpha*tmp_y).  thr holds 10 ints per partition in
// thr_info_t order (vbase, rbase, sbase, last_row, overflow_row, start_vec,
// end_vec, panel_id, merge_start, merge_end); scan masks take vec/8 bytes.
template <class T>
void vhcc_spmv_refsem(int64_t m, int vec, int nparts, int64_t pad_rows, const int* thr,
                      const uint32_t* veceor, const uint8_t* masks, const uint32_t* row_arr,
                      const uint32_t* col, const T* val, const T* x, T alpha, T beta, T* y) {
  int nm = 0;
  while ((1 << nm) < vec) ++nm;
  const int mb = vec / 8;
  auto bit = [&](int64_t mask_idx, int j) { return (masks[mask_idx * mb + j / 8] >> (j % 8)) & 1; };
  std::vector<T> tmp((size_t)pad_rows, T(0)), res((size_t)vec), sh((size_t)vec);
  for (int t = 0; t < nparts; ++t) {
    const int* I = thr + 10 * t;
    int64_t vidx = I[0], sidx = (int64_t)I[0] * (nm + 1), ridx = I[1], e = (int64_t)I[5] * vec;
    uint32_t eor = veceor[vidx++];
    std::fill(res.begin(), res.end(), T(0));
    for (int64_t v = I[5]; v < I[6]; ++v) {
      for (int j = 0; j < vec; ++j) res[(size_t)j] = val[e + j] * x[col[e + j]] + res[(size_t)j];
      e += vec;
      if ((uint32_t)v == eor) {
        for (int s = 0; s < nm; ++s, ++sidx) {
          const int d = 1 << s;
          for (int j = 0; j < vec; ++j) sh[(size_t)j] = j >= d ? res[(size_t)(j - d)] : T(0);
          for (int j = 0; j < vec; ++j)
            if (bit(sidx, j)) res[(size_t)j] = res[(size_t)j] + sh[(size_t)j];
        }
        const int64_t wr = sidx++;
        for (int j = 0; j < vec; ++j)
          if (bit(wr, j)) tmp[row_arr[ridx++]] += res[(size_t)j];
        eor = veceor[vidx++];
        const bool carry = !bit(wr, vec - 1);
        const T last = res[(size_t)vec - 1];
        std::fill(res.begin(), res.end(), T(0));
        if (carry) res[0] = last;
      } else {
        T s = 0;
        for (int j = 0; j < vec; ++j) s += res[(size_t)j];
        std::fill(res.begin(), res.end(), T(0));
        res[0] = s;
      }
    }
  }
  for (int t = 0; t < nparts; ++t) tmp[(size_t)thr[10 * t + 3]] += tmp[(size_t)thr[10 * t + 4]];
  for (int64_t i = 0; i < m; ++i) y[i] = y[i] * beta + alpha * tmp[(size_t)i];
}

// reference src/rvv/spmv_mtx_vnec.cpp:42-71 (MergePathDivide).
void merge_path_divide(int diagonal, const int* a, int a_len, int b_len, int* x, int* y) {
  int x_min = std::max(diagonal - b_len, 0);
  int x_max = std::min(diagonal, a_len);
  while (x_min < x_max) {
    const int pivot = (x_min + x_max) >> 1;
    if (a[pivot] <= diagonal - pivot - 1) x_min = pivot + 1;  // b[k] = k
    else x_max = pivot;
  }
  *x = std::min(x_min, a_len);
  *y = diagonal - x_min;
}

// reference include/VNEC.h:78-190 (IRD_VNEC), one thread walking the
// partitions in order (the reference's shared group_index is then
// race-free): merge-path coordinates, per-partition column compaction
// through the IDX_OFFSET prefix of not-null flags, lane groups.
float vnec_ird(int m, int n, int nnz, const int* rst, const int* col, int nthreads, int nlanes,
               int* coords) {
  const int items = nnz + m;
  const int per = (items + nthreads - 1) / nthreads;
  std::vector<int> sx(nthreads), sy(nthreads), ex(nthreads), ey(nthreads);
  for (int t = 0; t < nthreads; ++t) {
    const int d0 = std::min(per * t, items), d1 = std::min(d0 + per, items);
    merge_path_divide(d0, rst + 1, m, nnz, &sx[t], &sy[t]);
    merge_path_divide(d1, rst + 1, m, nnz, &ex[t], &ey[t]);
    if (coords) {
      coords[4 * t] = sx[t];
      coords[4 * t + 1] = sy[t];
      coords[4 * t + 2] = ex[t];
      coords[4 * t + 3] = ey[t];
    }
  }
  std::vector<int> ecr((size_t)nnz + 10, 0);
  std::vector<int> flag(n), off(n);
  for (int t = 0; t < nthreads; ++t) {
    for (int c = 0; c < n; ++c) flag[c] = 1;
    for (int j = sy[t]; j < ey[t]; ++j) flag[col[j]] = 0;
    off[0] = flag[0];
    for (int c = 1; c < n; ++c) off[c] = off[c - 1] + flag[c];
    for (int j = sy[t]; j < ey[t]; ++j) ecr[j] = col[j] - off[col[j]];
  }
  long long group = 0;
  for (int t = 0; t < nthreads; ++t) {
    for (int i = sx[t]; i < ex[t]; ++i) {
      const int start = rst[i] > sy[t] ? rst[i] : sy[t];
      const int len = rst[i + 1] - start;
      int first = ecr[start];
      for (int j = 1; j < len; j++)
        if (ecr[start + j] - first >= nlanes) {
          ++group;
          first = ecr[start + j];
        }
      if (len != 0) ++group;
    }
  }
  return (float)nnz / (float)(group * nlanes);
}

}  // namespace

extern "C" {

// IRD of reference include/VNEC.h:78-190 and the merge-path coordinates
// (start.x, start.y, end.x, end.y per partition) into coords (or NULL).
float orc_vnec_ird(int m, int n, int nnz, const int* rst, const int* col, int nthreads, int nlanes,
                   int* coords) {
  return vnec_ird(m, n, nnz, rst, col, nthreads, nlanes, coords);
}

int orc_vhcc_spmv_f64(int64_t m, int vec, int nparts, int64_t pad_rows, const int* thr,
                      const uint32_t* veceor, const uint8_t* masks, const uint32_t* row_arr,
                      const uint32_t* col, const double* val, const double* x, double alpha,
                      double beta, double* y) {
  vhcc_spmv_refsem<double>(m, vec, nparts, pad_rows, thr, veceor, masks, row_arr, col, val, x,
                           alpha, beta, y);
  return 0;
}
int orc_vhcc_spmv_f32(int64_t m, int vec, int nparts, int64_t pad_rows, const int* thr,
                      const uint32_t* veceor, const uint8_t* masks, const uint32_t* row_arr,
                      const uint32_t* col, const float* val, const float* x, float alpha,
                      float beta, float* y) {
  vhcc_spmv_refsem<float>(m, vec, nparts, pad_rows, thr, veceor, masks, row_arr, col, val, x,
                          alpha, beta, y);
  return 0;
}

void* orc_csr5_build_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                         const double* val, int omega, int sigma) {
  return csr5_build<double>(m, n, nnz, rst, col, val, omega, sigma);
}
void* orc_csr5_build_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                         const float* val, int omega, int sigma) {
  return csr5_build<float>(m, n, nnz, rst, col, val, omega, sigma);
}
// Reference-semantics CSR5 multiply over any CSR5 structure held in an Obj
// (the oracle's own or one copied in by orc_obj_from_arrays).
int orc_csr5_spmv_f64(void* h, const double* x, double alpha, double beta, double* y) {
  const Obj* o = (const Obj*)h;
  auto S = [&](const char* k) { return (int64_t)o->scalars.at(k); };
  csr5_spmv_refsem<double>(S("_m"), S("_p"), (int)S("_omega"), (int)S("_sigma"),
                           (int)S("_bit_y_offset"), (int)S("_bit_seg_offset"),
                           (int)S("_tile_desc_len"), (int)S("_tail_tile_start"),
                           o->get<int>("_csr_row_pointer"), o->get<uint32_t>("_csr_column_index"),
                           o->get<double>("_csr_value"), o->get<uint32_t>("_csr5_tile_pointer"),
                           o->get<uint32_t>("_csr5_tile_descriptor"),
                           o->get<uint32_t>("_csr5_tile_empty_offset_pointer"),
                           o->get<uint32_t>("_csr5_tile_empty_offset"), x, alpha, beta, y);
  return 0;
}
int orc_csr5_spmv_f32(void* h, const float* x, float alpha, float beta, float* y) {
  const Obj* o = (const Obj*)h;
  auto S = [&](const char* k) { return (int64_t)o->scalars.at(k); };
  csr5_spmv_refsem<float>(S("_m"), S("_p"), (int)S("_omega"), (int)S("_sigma"),
                          (int)S("_bit_y_offset"), (int)S("_bit_seg_offset"),
                          (int)S("_tile_desc_len"), (int)S("_tail_tile_start"),
                          o->get<int>("_csr_row_pointer"), o->get<uint32_t>("_csr_column_index"),
                          o->get<float>("_csr_value"), o->get<uint32_t>("_csr5_tile_pointer"),
                          o->get<uint32_t>("_csr5_tile_descriptor"),
                          o->get<uint32_t>("_csr5_tile_empty_offset_pointer"),
                          o->get<uint32_t>("_csr5_tile_empty_offset"), x, alpha, beta, y);
  return 0;
}

// An empty structure to be filled with arrays and scalars from outside (so
// the library's layouts can be run through the restated kernels).
void* orc_obj_new(void) { return new Obj; }
int orc_obj_set_array(void* h, const char* name, const void* data, int64_t count, int elem) {
  Obj* o = (Obj*)h;
  std::vector<char> b((size_t)(count * elem));
  if (count) std::memcpy(b.data(), data, b.size());
  o->arrays[name] = {std::move(b), elem};
  return 0;
}
int orc_obj_set_scalar(void* h, const char* name, double v) {
  ((Obj*)h)->scalars[name] = v;
  return 0;
}

int orc_csr_spmv_f64(int64_t m, const int* rst, const int* col, const double* val, const double* b,
                     double alpha, double beta, double* y) {
  csr_f64(m, rst, col, val, b, alpha, beta, y);
  return 0;
}

int orc_csr_spmv_f32(int64_t m, const int* rst, const int* col, const float* val, const float* b,
                     float alpha, float beta, float* y) {
  csr_f32(m, rst, col, val, b, alpha, beta, y);
  return 0;
}

// Per-row sum |a_ij| * |x_j| (the scale of the error bound), in double.
int orc_csr_absdot(int64_t m, const int* rst, const int* col, const double* val, const double* b,
                   double* out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < m; i++) {
    double t = 0.0;
    for (int j = rst[i]; j < rst[i + 1]; j++) t += std::fabs(val[j]) * std::fabs(b[col[j]]);
    out[i] = t;
  }
  return 0;
}

// Row dot products accumulated in double-double (TwoProd via fma + TwoSum),
// rounded once: a reference accurate to ~1e-30 relative for the error
// statistics of reduced-precision results.
int orc_csr_dot_dd(int64_t m, const int* rst, const int* col, const double* val, const double* b,
                   double* out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < m; i++) {
    double hi = 0.0, lo = 0.0;
    for (int j = rst[i]; j < rst[i + 1]; j++) {
      const double p = val[j] * b[col[j]];
      const double pe = std::fma(val[j], b[col[j]], -p);
      const double s = hi + p;
      const double z = s - hi;
      const double e = (hi - (s - z)) + (p - z);
      hi = s;
      lo += e + pe;
    }
    out[i] = hi + lo;
  }
  return 0;
}

void* orc_hcsr_convert_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                           const double* val, int64_t R, int64_t C, double p, int* err) {
  return hcsr_convert<double>(m, n, nnz, rst, col, val, R, C, p, err);
}
void* orc_hcsr_convert_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                           const float* val, int64_t R, int64_t C, double p, int* err) {
  return hcsr_convert<float>(m, n, nnz, rst, col, val, R, C, p, err);
}
int orc_hcsr_spmv_f64(void* h, const double* b, double alpha, double beta, double* y) {
  hcsr_spmv<double>((Obj*)h, b, alpha, beta, y);
  return 0;
}
int orc_hcsr_spmv_f32(void* h, const float* b, float alpha, float beta, float* y) {
  hcsr_spmv<float>((Obj*)h, b, alpha, beta, y);
  return 0;
}

void* orc_sell_build_f64(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                         const double* val, int C, int sigma) {
  return sell_build<double>(m, n, nnz, rst, col, val, C, sigma);
}
void* orc_sell_build_f32(int64_t m, int64_t n, int64_t nnz, const int* rst, const int* col,
                         const float* val, int C, int sigma) {
  return sell_build<float>(m, n, nnz, rst, col, val, C, sigma);
}
int orc_sell_spmv_f64(int64_t m, int64_t nb, const int* Cl, const int* Cs, const int* Bs,
                      const int* SBs, const uint32_t* Col, const uint32_t* Perm, const double* Val,
                      const double* b, double alpha, double beta, double* y) {
  sell_spmv<double>(m, nb, Cl, Cs, Bs, SBs, Col, Perm, Val, b, alpha, beta, y);
  return 0;
}
int orc_sell_spmv_f32(int64_t m, int64_t nb, const int* Cl, const int* Cs, const int* Bs,
                      const int* SBs, const uint32_t* Col, const uint32_t* Perm, const float* Val,
                      const float* b, float alpha, float beta, float* y) {
  sell_spmv<float>(m, nb, Cl, Cs, Bs, SBs, Col, Perm, Val, b, alpha, beta, y);
  return 0;
}

// Generic access to a converted structure.
int orc_obj_array(void* h, const char* name, const void** data, int64_t* count, int* elem) {
  Obj* o = (Obj*)h;
  auto it = o->arrays.find(name);
  if (it == o->arrays.end()) return -23;
  *data = it->second.first.data();
  *elem = it->second.second;
  *count = (int64_t)it->second.first.size() / it->second.second;
  return 0;
}
int orc_obj_scalar(void* h, const char* name, double* v) {
  Obj* o = (Obj*)h;
  auto it = o->scalars.find(name);
  if (it == o->scalars.end()) return -23;
  *v = it->second;
  return 0;
}
void orc_obj_free(void* h) { delete (Obj*)h; }

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

}  // extern "C"
