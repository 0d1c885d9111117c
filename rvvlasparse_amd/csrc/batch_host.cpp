// batch_host.cpp -- batched uneven SpMM in one persistent launch, with an
// autotuner that picks each matrix's item kernel and its parameters.
//
// The autotuner keys matrices by their row-length histogram (fractions of
// rows in power-of-two length buckets, rounded to 1/20, the max/mean
// ratio bucket and the padded right-hand-side count) and times the
// candidate item kernels on the first matrix of each key, on the device,
// with HIP events; the winner is cached for the process (and, with the
// parameter cache=<path>, in a text file of "key fmt param" lines).
// Candidates: CSR rows with the slot count chosen from the mean row length,
// with a quarter and with four times as many slots, and SELL-64 with sigma
// 1 or 4096.
//
// Work items: after all matrices are laid out, each matrix is cut into runs
// of rows (CSR) or slices (SELL) of about equal cost, about 128 items per
// wavefront of the persistent grid (parameter wave_items): few enough that
// the queue costs little in atomics, enough that the waves of an XCD work
// on about one matrix at a time (whose X then stays in its L2 when it fits)
// and that the last items finish together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "devmat.h"
#include "internal.h"

using namespace SparseMatrixLib;
using rvvla::BatchItem;
using rvvla::BatchMat;
using rvvla::Error;

namespace {

struct Choice {
  int fmt = 0;    // 0 = CSR rows, 1 = SELL-64 slices
  int param = 0;  // CSR: slots per row (0 = from the mean row length); SELL: sigma
};

constexpr int kBatchGrid = 2048;  // persistent workgroups (4 wavefronts each)

std::mutex g_tune_mu;
std::map<std::string, Choice> g_tune_cache;

}  // namespace

struct rvvla_batch_s {
  int dtype = RVVLA_F32;
  int count = 0;
  int grid = 1;
  std::vector<BatchMat> mats;
  std::vector<BatchItem> items;
  std::vector<std::string> keys;
  std::vector<void*> allocs;
  rvvla::BatchArgs args{};
  const void* val = nullptr;
  const void* sval = nullptr;
  void* X = nullptr;
  void* Y = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_alpha = 0, graph_beta = 0;

  ~rvvla_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : allocs) (void)hipFree(p);
  }
  void* put(const void* host, size_t bytes) {
    void* d = nullptr;
    int rc = rvvla::hip_check(hipMalloc(&d, bytes ? bytes : 8), "hipMalloc");
    if (rc) throw Error(rc, rvvla::last_error());
    allocs.push_back(d);
    if (host && bytes) {
      rc = rvvla::hip_check(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice), "hipMemcpy");
      if (rc) throw Error(rc, rvvla::last_error());
    }
    return d;
  }
};

namespace {

template <class V>
struct HostMat {
  int64_t m, n;
  const int* rst;
  const int* col;
  const V* val;
  int k;
};

// Right-hand sides per lane: the largest power of two <= vmax dividing k.
static int vec_width(int k, int vmax) {
  int v = 1;
  while (v * 2 <= vmax && k % (v * 2) == 0) v *= 2;
  return v;
}

static int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Appends matrices under their choices, then cuts the work items.
template <class V>
struct Builder {
  std::vector<int32_t> rst, col, cl, cs, scol, perm;
  std::vector<V> val, sval;
  std::vector<BatchMat> mats;
  std::vector<BatchItem> items;
  int64_t x_elems = 0, y_elems = 0;
  int wave_items = 128;  // work items per wavefront of the persistent grid

  void add(const HostMat<V>& h, const Choice& c) {
    BatchMat M;
    M.m = (int32_t)h.m;
    M.n = (int32_t)h.n;
    M.k = h.k;
    M.fmt = c.fmt;
    // rows of X_i and Y_i padded to whole vector loads (16 bytes) when k is
    // not a multiple of them: one lane then reads 4 (FP64: 2) right-hand
    // sides at once instead of one
    const int vmax = sizeof(V) == 8 ? 2 : 4;
    M.ld = h.k <= 2 ? h.k : (h.k + vmax - 1) / vmax * vmax;
    M.vw = vec_width(M.ld, vmax);
    M.lpe = pow2_ceil(M.ld / M.vw);
    // X and Y blocks start on 4-element boundaries (vector loads)
    x_elems = (x_elems + 3) & ~int64_t(3);
    y_elems = (y_elems + 3) & ~int64_t(3);
    M.x_off = x_elems;
    M.y_off = y_elems;
    x_elems += h.n * M.ld;
    y_elems += h.m * M.ld;
    if (c.fmt == 0) {
      const int64_t nnz = h.rst[h.m];
      int slots = c.param;
      if (slots <= 0) {
        const double mean = h.m ? (double)nnz / (double)h.m : 1.0;
        slots = 1;
        while (slots * 2 <= 64 && slots * 2 * 4 <= mean) slots *= 2;  // ~4 entries per lane
      }
      slots = std::max(1, std::min(pow2_ceil(slots), 64 / M.lpe));
      M.gw = M.lpe * slots;
      M.rst_off = (int64_t)rst.size();
      M.ent_off = (int64_t)col.size();
      rst.insert(rst.end(), h.rst, h.rst + h.m + 1);
      col.insert(col.end(), h.col, h.col + nnz);
      val.insert(val.end(), h.val, h.val + nnz);
    } else {
      spMtxCRS<V> a((size_t)h.m, (size_t)h.n, (size_t)h.rst[h.m], h.rst, h.col, h.val);
      spMtxSELL_C_Sigma<V> s;
      rvvla::build_sell<V>(s, a, 64, c.param);
      M.sl_off = (int64_t)cl.size();
      M.slot_off = (int64_t)scol.size();
      M.row_off = (int64_t)perm.size();
      cl.insert(cl.end(), s.vCl.begin(), s.vCl.end());
      cs.insert(cs.end(), s.vCs.begin(), s.vCs.end());
      scol.insert(scol.end(), s.vCol.begin(), s.vCol.end());
      sval.insert(sval.end(), s.vVal.begin(), s.vVal.end());
      perm.insert(perm.end(), s.vPerm.begin(), s.vPerm.end());
    }
    mats.push_back(M);
  }

  // Cost units (wavefront load steps): a CSR row costs len * lpe / 64 plus a
  // fixed share for its pointers and reduction, a SELL slice width * k.
  double unit_cost(const BatchMat& M, int64_t u) const {
    if (M.fmt == 0) {
      const int64_t len = rst[(size_t)(M.rst_off + u + 1)] - rst[(size_t)(M.rst_off + u)];
      return ((double)len * M.lpe + 4.0 * M.gw) / 64.0;
    }
    return (double)cl[(size_t)(M.sl_off + u)] * M.k + 2.0;
  }

  // Cuts every matrix into items of about `target` cost and packs the
  // matrices into `nq` queues, longest processing time first (each matrix
  // to the least loaded queue); a queue lists its matrices' items in order.
  void cut_items(int waves, int nq, std::vector<int64_t>& qbase) {
    items.clear();
    std::vector<double> cost(mats.size(), 0.0);
    double total = 0;
    for (size_t i = 0; i < mats.size(); ++i) {
      const BatchMat& M = mats[i];
      const int64_t units = M.fmt == 0 ? M.m : (M.m + 63) / 64;
      for (int64_t u = 0; u < units; ++u) cost[i] += unit_cost(M, u);
      total += cost[i];
    }
    const double target = std::max(16.0, total / ((double)wave_items * std::max(1, waves)));
    std::vector<size_t> order(mats.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<double> load((size_t)nq, 0.0);
    std::vector<std::vector<size_t>> queue((size_t)nq);
    for (size_t i : order) {
      const size_t q = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
      load[q] += cost[i];
      queue[q].push_back(i);
    }
    qbase.assign((size_t)nq + 1, 0);
    for (int q = 0; q < nq; ++q) {
      for (size_t i : queue[(size_t)q]) {
        const BatchMat& M = mats[i];
        const int64_t units = M.fmt == 0 ? M.m : (M.m + 63) / 64;
        const int64_t step = M.fmt == 0 ? 64 / M.gw : 1;  // rows per wavefront step
        int64_t u0 = 0;
        double acc = 0;
        for (int64_t u = 0; u < units; ++u) {
          acc += unit_cost(M, u);
          const bool cut = acc >= target && (u + 1 - u0) % step == 0;
          if (cut || u + 1 == units) {
            items.push_back(BatchItem{(int32_t)i, (int32_t)u0, (int32_t)(u + 1 - u0), 0});
            u0 = u + 1;
            acc = 0;
          }
        }
      }
      qbase[(size_t)q + 1] = (int64_t)items.size();
    }
  }

  // move everything to the device
  void upload(rvvla_batch_s& B, int dtype) {
    int xcc = 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess)
      xcc = 1;
    const int nq = std::max(1, std::min(xcc, rvvla::kMaxQueues));
    std::vector<int64_t> qbase;
    cut_items(kBatchGrid * (rvvla::kWG / 64), nq, qbase);
    B.args.nqueues = nq;
    for (int q = 0; q <= nq; ++q) B.args.qbase[q] = qbase[(size_t)q];
    B.mats = mats;
    B.items = items;
    B.args.mats = (const BatchMat*)B.put(mats.data(), mats.size() * sizeof(BatchMat));
    B.args.items = (const BatchItem*)B.put(items.data(), items.size() * sizeof(BatchItem));
    B.args.nitems = (int64_t)items.size();
    B.args.counter = (int*)B.put(nullptr, (size_t)rvvla::kMaxQueues * rvvla::kQueueStride * 4);
    B.args.rst = (const int32_t*)B.put(rst.data(), rst.size() * 4);
    B.args.col = (const int32_t*)B.put(col.data(), col.size() * 4);
    B.args.cl = (const int32_t*)B.put(cl.data(), cl.size() * 4);
    B.args.cs = (const int32_t*)B.put(cs.data(), cs.size() * 4);
    B.args.scol = (const int32_t*)B.put(scol.data(), scol.size() * 4);
    B.args.perm = (const int32_t*)B.put(perm.data(), perm.size() * 4);
    B.val = B.put(val.data(), val.size() * sizeof(V));
    B.sval = B.put(sval.data(), sval.size() * sizeof(V));
    B.X = B.put(nullptr, (size_t)x_elems * sizeof(V));
    B.Y = B.put(nullptr, (size_t)y_elems * sizeof(V));
    int rc = rvvla::hip_check(hipMemset(B.X, 0, (size_t)x_elems * sizeof(V)), "clear X");
    if (!rc) rc = rvvla::hip_check(hipMemset(B.Y, 0, (size_t)y_elems * sizeof(V)), "clear Y");
    if (rc) throw Error(rc, rvvla::last_error());
    B.dtype = dtype;
    B.count = (int)mats.size();
    B.grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)items.size() + 3) / 4, kBatchGrid));
  }
};

template <class V>
double time_choice(const HostMat<V>& h, const Choice& c, int dtype) {
  Builder<V> b;
  b.add(h, c);
  rvvla_batch_s B;
  b.upload(B, dtype);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int w = 0; w < 2; ++w)
    rvvla::launch_batch(dtype, B.args, B.val, B.sval, B.X, B.Y, 1.0, 0.0, B.grid, nullptr);
  (void)hipEventRecord(e0, nullptr);
  const int reps = 5;
  for (int w = 0; w < reps; ++w)
    rvvla::launch_batch(dtype, B.args, B.val, B.sval, B.X, B.Y, 1.0, 0.0, B.grid, nullptr);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

void load_cache(const std::string& path) {
  std::ifstream f(path);
  std::string key;
  Choice c;
  while (f >> key >> c.fmt >> c.param) g_tune_cache[key] = c;
}

void save_cache(const std::string& path) {
  std::ofstream f(path);
  for (const auto& kv : g_tune_cache) f << kv.first << " " << kv.second.fmt << " " << kv.second.param << "\n";
}

template <class V>
rvvla_batch_t create(int dtype, int count, const int64_t* m, const int64_t* n,
                     const int32_t* const* rst, const int32_t* const* col,
                     const void* const* val, const int* k, const convertParams& P,
                     const std::string& cache_path) {
  const int tune = rvvla::param_int(P, "tune", 1);
  const int force_fmt = rvvla::param_int(P, "format", -1);
  std::vector<HostMat<V>> hm((size_t)count);
  for (int i = 0; i < count; ++i) {
    if (m[i] < 0 || n[i] < 1 || k[i] < 1 || k[i] > 16 || !rst[i] || (rst[i][m[i]] > 0 && (!col[i] || !val[i])))
      throw Error(RVVLA_ERR_INVALID_ARG, "batch: bad matrix " + std::to_string(i));
    rvvla::check_csr_public(m[i], n[i], rst[i][m[i]], rst[i], col[i]);
    hm[(size_t)i] = HostMat<V>{m[i], n[i], rst[i], col[i], (const V*)val[i], k[i]};
  }
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (!cache_path.empty()) load_cache(cache_path);
  auto B = std::make_unique<rvvla_batch_s>();
  Builder<V> b;
  b.wave_items = rvvla::param_int(P, "wave_items", 128);
  if (b.wave_items < 1 || b.wave_items > 1024)
    throw Error(RVVLA_ERR_INVALID_ARG, "batch: wave_items must be in [1, 1024]");
  for (int i = 0; i < count; ++i) {
    const HostMat<V>& h = hm[(size_t)i];
    const std::string key = rvvla::row_length_key(h.m, h.rst, h.k);
    Choice c;
    if (force_fmt >= 0) {
      c.fmt = force_fmt;
      c.param = rvvla::param_int(P, "param", force_fmt == 0 ? 0 : 1);
    } else if (g_tune_cache.count(key)) {
      c = g_tune_cache[key];
    } else if (tune) {
      // CSR slot counts around the one the mean row length suggests
      const int64_t nnz = h.rst[h.m];
      const double mean = h.m ? (double)nnz / (double)h.m : 1.0;
      int s0 = 1;
      while (s0 * 2 <= 64 && s0 * 2 * 4 <= mean) s0 *= 2;
      const Choice cands[] = {{0, s0}, {0, std::max(1, s0 / 4)}, {0, std::min(64, s0 * 4)},
                              {1, 1}, {1, 4096}};
      double best = 1e30;
      for (const Choice& cand : cands) {
        const double t = time_choice<V>(h, cand, dtype);
        if (t < best) {
          best = t;
          c = cand;
        }
      }
      g_tune_cache[key] = c;
    }
    B->keys.push_back(key);
    b.add(h, c);
  }
  b.upload(*B, dtype);
  if (!cache_path.empty()) save_cache(cache_path);
  return B.release();
}

}  // namespace

extern "C" {

int rvvla_batch_create(int dtype, int count, const int64_t* m, const int64_t* n,
                       const int32_t* const* rst, const int32_t* const* col,
                       const void* const* val, const int* k, const char* params,
                       rvvla_batch_t* out) {
  try {
    if (!out || count < 0 || (count > 0 && (!m || !n || !rst || !col || !val || !k)))
      return RVVLA_ERR_INVALID_ARG;
    *out = nullptr;
    if (dtype != RVVLA_F64 && dtype != RVVLA_F32) return RVVLA_ERR_VALUE_TYPE;
    // a cache path is not a number: take it out verbatim before parsing the rest
    std::string cache, rest = params ? params : "";
    const size_t p = rest.find("cache=");
    if (p != std::string::npos) {
      const size_t e = rest.find_first_of(",;", p);
      cache = rest.substr(p + 6, e == std::string::npos ? std::string::npos : e - p - 6);
      rest.erase(p, e == std::string::npos ? std::string::npos : e - p + 1);
    }
    const convertParams P = rvvla::parse_params(rest.c_str());
    *out = dtype == RVVLA_F64
               ? create<double>(dtype, count, m, n, rst, col, val, k, P, cache)
               : create<float>(dtype, count, m, n, rst, col, val, k, P, cache);
    return RVVLA_OK;
  } catch (const Error& e) {
    rvvla::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    rvvla::set_last_error(e.what());
    return RVVLA_ERR_LOGIC;
  }
}

int rvvla_batch_block(rvvla_batch_t B, int i, void** x, void** y, int64_t* ld, int* format) {
  if (!B || i < 0 || i >= B->count) return RVVLA_ERR_INVALID_ARG;
  const size_t es = B->dtype == RVVLA_F64 ? 8 : 4;
  const BatchMat& M = B->mats[(size_t)i];
  if (x) *x = (char*)B->X + M.x_off * es;
  if (y) *y = (char*)B->Y + M.y_off * es;
  if (ld) *ld = M.ld;
  if (format) *format = M.fmt == 0 ? RVVLA_FMT_CSR : RVVLA_FMT_SELL;
  return RVVLA_OK;
}

int rvvla_batch_run(rvvla_batch_t B, double alpha, double beta, void* stream) {
  if (!B) return RVVLA_ERR_INVALID_ARG;
  return rvvla::launch_batch(B->dtype, B->args, B->val, B->sval, B->X, B->Y, alpha, beta, B->grid,
                             (hipStream_t)stream);
}

int rvvla_batch_capture(rvvla_batch_t B, double alpha, double beta) {
  if (!B) return RVVLA_ERR_INVALID_ARG;
  hipStream_t s;
  int rc = rvvla::hip_check(hipStreamCreate(&s), "hipStreamCreate");
  if (rc) return rc;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  rc = rvvla::hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  if (!rc) {
    const int rc2 = rvvla_batch_run(B, alpha, beta, s);
    rc = rvvla::hip_check(hipStreamEndCapture(s, &B->graph), "end capture");
    if (!rc) rc = rc2;
  }
  if (!rc) rc = rvvla::hip_check(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0), "instantiate");
  (void)hipStreamDestroy(s);
  B->graph_alpha = alpha;
  B->graph_beta = beta;
  return rc;
}

int rvvla_batch_replay(rvvla_batch_t B, void* stream) {
  if (!B || !B->exec) return RVVLA_ERR_INVALID_ARG;
  return rvvla::hip_check(hipGraphLaunch(B->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int rvvla_batch_free(rvvla_batch_t B) {
  delete B;
  return RVVLA_OK;
}

int rvvla_batch_tune_key(rvvla_batch_t B, int i, char* buf, int len) {
  if (!B || i < 0 || i >= B->count || !buf || len < 1) return RVVLA_ERR_INVALID_ARG;
  std::snprintf(buf, (size_t)len, "%s", B->keys[(size_t)i].c_str());
  return RVVLA_OK;
}

}  // extern "C"
