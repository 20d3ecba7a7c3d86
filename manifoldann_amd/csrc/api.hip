// C ABI (include/mann.h): argument checks with the reference's exception
// classes, host/device pointer staging, and dispatch to the pipelines.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <vector>

#include "../../include/mann.h"
#include "core.h"
#include "hnsw.h"
#include "fp8.h"
#include "ivf.h"
#include "mlp.h"

namespace mann {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

template <class F>
static int guarded(F&& f) {
  try {
    g_err.clear();
    tuning_refresh();
    f();
    return OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return ERR_ALLOC;
  } catch (const std::exception& e) {
    set_error(e.what());
    return ERR_DEVICE;
  }
}

// Device destination for results; copied back to host memory if needed.
template <class T>
struct OutBuf {
  T* user;
  T* dev;
  size_t count;
  bool staged;
  OutBuf(T* u, size_t cnt, const char* tag) : user(u), dev(u), count(cnt), staged(false) {
    if (u && !is_device_ptr(u)) {
      dev = ctx().buf(tag).as<T>(cnt > 0 ? cnt : 1);
      staged = true;
    }
  }
  void finish() {
    if (staged && user && count) {
      hipStream_t st = ctx().stream;
      MANN_HIP(hipMemcpyAsync(user, dev, sizeof(T) * count, hipMemcpyDeviceToHost, st));
      MANN_HIP(hipStreamSynchronize(st));
    }
  }
};

static void check_metric(int metric) {
  MANN_REQUIRE(metric == EUCLIDEAN || metric == SQEUCLIDEAN || metric == COSINE, ERR_ARGUMENT,
               "unknown metric");
}

}  // namespace mann

using namespace mann;

extern "C" {

int mann_version(void) { return 10000; }

const char* mann_last_error(void) { return g_err.c_str(); }

int mann_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int mann_dataset_create(const float* data, int64_t d, int64_t n, int64_t ld, int metric,
                        mann_dataset_t* out) {
  return guarded([&] {
    MANN_REQUIRE(out != nullptr, ERR_NULL, "null output handle");
    MANN_REQUIRE(data != nullptr, ERR_NULL, "null data pointer");
    check_metric(metric);
    *out = create_dataset(data, d, n, ld, metric);
  });
}

int mann_dataset_create_f16(const void* data, int64_t d, int64_t n, int64_t ld, int metric,
                            mann_dataset_t* out) {
  return guarded([&] {
    MANN_REQUIRE(out != nullptr, ERR_NULL, "null output handle");
    MANN_REQUIRE(data != nullptr, ERR_NULL, "null data pointer");
    MANN_REQUIRE(d > 0 && n > 0, ERR_ARGUMENT, "Dataset must contain at least one point");
    MANN_REQUIRE(ld >= d, ERR_ARGUMENT, "leading dimension smaller than d");
    check_metric(metric);
    ThreadCtx& c = ctx();
    hipStream_t st = c.stream;
    const size_t elems = (size_t)(ld * (n - 1) + d);
    struct Tmp {  // freed on every exit (hipFree waits for the device)
      void* p = nullptr;
      ~Tmp() {
        if (p) (void)hipFree(p);
      }
    } staged16, up;
    const void* src = data;
    if (!is_device_ptr(data)) {
      MANN_HIP(hipMalloc(&staged16.p, elems * 2));
      MANN_HIP(hipMemcpyAsync(staged16.p, data, elems * 2, hipMemcpyHostToDevice, st));
      src = staged16.p;
    }
    // exact upcast to FP32 rows of stride d, then the FP32 preparation
    MANN_HIP(hipMalloc(&up.p, sizeof(float) * (size_t)n * d));
    f16_rows_to_f32(src, ld, (float*)up.p, d, (int)d, n, st);
    *out = create_dataset((const float*)up.p, d, n, d, metric);
  });
}

int mann_dataset_destroy(mann_dataset_t ds) {
  return guarded([&] { delete ds; });
}

int mann_dataset_info(mann_dataset_t ds, int64_t* d, int64_t* n) {
  return guarded([&] {
    MANN_REQUIRE(ds != nullptr, ERR_NULL, "null dataset");
    if (d) *d = ds->d;
    if (n) *n = ds->n;
  });
}

int mann_dataset_set_first_pass(mann_dataset_t ds, int kind) {
  return guarded([&] {
    MANN_REQUIRE(ds != nullptr, ERR_NULL, "null dataset");
    MANN_HIP(hipSetDevice(ds->device));
    dataset_set_first_pass(*ds, kind);
  });
}

int mann_bf_knn(mann_dataset_t base, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                int64_t k, int metric, int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(base != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    // generic batch query checks rows first, even for an empty batch (ann_index.jl:86-87)
    MANN_REQUIRE(d == base->d, ERR_DIMENSION,
                 "Expected queries with " + std::to_string(base->d) + " rows, got " +
                     std::to_string(d));
    MANN_REQUIRE(nq >= 0, ERR_ARGUMENT, "negative query count");
    if (nq == 0 || k <= 0) {
      // k <= 0 -> empty neighbour lists (bruteforce.jl:56)
      last_stats() = SearchStats();
      return;
    }
    MANN_REQUIRE(queries && out_ids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(ldq >= d, ERR_ARGUMENT, "leading dimension smaller than d");
    MANN_REQUIRE((metric == COSINE) == (base->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    MANN_REQUIRE(nq < (1 << 30), ERR_UNSUPPORTED, "too many queries in one call");
    QueryOps Q = prepare_queries(queries, d, nq, ldq, base->normalized, (int)base->d16,
                                 (int)base->kp, "bfq", base->center);
    OutBuf<int64_t> oi(out_ids, (size_t)(nq * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(nq * k), "out_dist");
    bf_search(*base, Q, (int)k, metric, 0, oi.dev, od.dev, nullptr);
    oi.finish();
    od.finish();
    timer_flush();
  });
}

// ---- allocation-free exact search ---------------------------------------------
//
// The scratch of mann_bf_knn_noalloc: one file-scope device array, which every
// GPU the library runs on holds until the process ends.  Host code never uses
// its name as an address: the address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept per device.
constexpr size_t kNoAllocScratch = (size_t)64 << 20;
constexpr size_t kNoAllocGuard = (size_t)64 << 10;  // never carved: the array's last 64 KiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

static constexpr int kMaxDevices = 64;
struct NoAllocDevice {
  std::once_flag once;
  char* ptr = nullptr;
  hipError_t rc = hipSuccess;
  std::mutex mu;  // calls on one device run one at a time (they share the array)
};
static NoAllocDevice g_noalloc[kMaxDevices];

static char* noalloc_scratch(int dev) {
  MANN_REQUIRE(dev >= 0 && dev < kMaxDevices, ERR_DEVICE, "device index out of range");
  NoAllocDevice& nd = g_noalloc[dev];
  std::call_once(nd.once, [&] {
    void* ptr = nullptr;
    nd.rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    nd.ptr = (char*)ptr;
  });
  MANN_REQUIRE(nd.rc == hipSuccess && nd.ptr != nullptr, ERR_DEVICE,
               std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(nd.rc));
  return nd.ptr;
}

// Upper bound of the scratch an exact search of nq queries carves before its
// exact fallback (which sizes its batches to what is left, down to one query).
static size_t noalloc_need(const mann_dataset& ds, int64_t nq, int64_t k) {
  const size_t nq_pad = (size_t)round_up(nq, 256);
  size_t b = 0;
  b += nq_pad * ds.d16 * 4 + nq_pad * ds.kp * 2 + 4 * nq_pad * 4;  // query operands
  if (ds.fold) b += nq_pad * ds.kp * 2;                              // folded operand
  b += (size_t)nq * select_cap_max() * 8 + (size_t)nq * 64 * 4;      // candidates, counts
  b += 3 * (size_t)nq * 4 + 32 + kStatSlots * 16;                    // thresholds, lists
  b += ((size_t)round_up(nq, 256) + 256) * 4;                        // padded threshold
  b += (size_t)256 * ds.m * 4;                                       // one sample chunk
  const size_t kr = (size_t)std::min<int64_t>(k, exact_kmax());
  b += (size_t)ceil_div(ds.n, exact_seg_rows()) * kr * 8;            // one fallback query
  return b + b / 8 + 40 * 512 + kNoAllocGuard;  // carve slack and alignment, end guard
}

int mann_bf_knn_noalloc(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                        int64_t ldq, int64_t k, int metric, int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(base != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    MANN_REQUIRE(d == base->d, ERR_DIMENSION,
                 "Expected queries with " + std::to_string(base->d) + " rows, got " +
                     std::to_string(d));
    MANN_REQUIRE(nq >= 0, ERR_ARGUMENT, "negative query count");
    if (nq == 0 || k <= 0) {
      last_stats() = SearchStats();
      return;
    }
    MANN_REQUIRE(queries && out_ids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(ldq >= d, ERR_ARGUMENT, "leading dimension smaller than d");
    MANN_REQUIRE((metric == COSINE) == (base->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    MANN_REQUIRE(is_device_ptr(queries) && is_device_ptr(out_ids) && is_device_ptr(out_dist),
                 ERR_ARGUMENT,
                 "the allocation-free search takes device queries and outputs only "
                 "(host buffers would need staging allocations)");
    // larger k takes the all-distances sort, whose scratch grows with n
    MANN_REQUIRE(k <= select_exmax() / 2, ERR_ARGUMENT,
                 "k = " + std::to_string(k) + " exceeds the allocation-free search's maximum of " +
                     std::to_string(select_exmax() / 2) + " neighbours per query");
    const size_t need = noalloc_need(*base, nq, k);
    MANN_REQUIRE(need <= kNoAllocScratch, ERR_ARGUMENT,
                 std::to_string(nq) + " queries need " + std::to_string(need) +
                     " bytes of scratch; the allocation-free search has " +
                     std::to_string(kNoAllocScratch) + ": split the query batch");
    ThreadCtx& c = ctx();
    char* sp = noalloc_scratch(c.device);
    std::lock_guard<std::mutex> lock(g_noalloc[c.device].mu);
    Arena arena{sp, kNoAllocScratch - kNoAllocGuard, 0};
    c.abufs.clear();
    c.arena = &arena;
    struct Restore {
      ThreadCtx& c;
      ~Restore() {
        c.arena = nullptr;
        c.abufs.clear();
      }
    } restore{c};
    QueryOps Q = prepare_queries(queries, d, nq, ldq, base->normalized, (int)base->d16,
                                 (int)base->kp, "bfq", base->center);
    bf_search(*base, Q, (int)k, metric, 0, out_ids, out_dist, nullptr);
  });
}

// ---- exact search captured in a HIP graph ---------------------------------------

struct mann_graph {
  int device = 0;
  int nq = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* arena_mem = nullptr;  // every scratch buffer of the captured search
  // (or, for a captured IVF search, its named scratch buffers)
  std::map<std::string, mann::DevBuf> owned;
  mann::CaptureInfo ci;
  ~mann_graph() {
    (void)hipSetDevice(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (arena_mem) (void)hipFree(arena_mem);
    for (auto& kv : owned) kv.second.release();
  }
};

int mann_bf_knn_graph_create(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                             int64_t ldq, int64_t k, int metric, int64_t* out_ids,
                             float* out_dist, mann_graph_t* out) {
  return guarded([&] {
    MANN_REQUIRE(out != nullptr, ERR_NULL, "null output handle");
    MANN_REQUIRE(base != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    MANN_REQUIRE(d == base->d, ERR_DIMENSION,
                 "Expected queries with " + std::to_string(base->d) + " rows, got " +
                     std::to_string(d));
    MANN_REQUIRE(nq > 0 && k > 0, ERR_ARGUMENT, "a captured search needs queries and k > 0");
    MANN_REQUIRE(k <= select_exmax() / 2, ERR_ARGUMENT,
                 "k = " + std::to_string(k) + " exceeds the GPU search's maximum of " +
                     std::to_string(select_exmax() / 2) + " neighbours per query");
    MANN_REQUIRE(queries && out_ids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(ldq >= d, ERR_ARGUMENT, "leading dimension smaller than d");
    MANN_REQUIRE((metric == COSINE) == (base->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    MANN_REQUIRE(is_device_ptr(queries) && is_device_ptr(out_ids) && is_device_ptr(out_dist),
                 ERR_ARGUMENT, "a captured search takes device queries and outputs only");
    ThreadCtx& c = ctx();
    auto g = std::make_unique<mann_graph>();
    g->device = c.device;
    g->nq = (int)nq;
    // the exact fallback inside the graph covers this many flagged queries
    const int F = (int)std::min<int64_t>(nq, 64);
    const size_t kr = (size_t)std::min<int64_t>(k, exact_kmax());
    const size_t per = (size_t)ceil_div(base->n, exact_seg_rows()) * kr * 8;
    const size_t need = noalloc_need(*base, nq, k) + (size_t)F * (per + 256) + ((size_t)1 << 20);
    MANN_HIP(hipMalloc(&g->arena_mem, need));
    Arena arena{(char*)g->arena_mem, need, 0};
    struct Restore {
      ThreadCtx& c;
      ~Restore() {
        c.arena = nullptr;
        c.capture = nullptr;
        c.abufs.clear();
      }
    } restore{c};
    auto run = [&] {
      c.abufs.clear();
      arena.used = 0;
      c.arena = &arena;
      QueryOps Q = prepare_queries(queries, d, nq, ldq, base->normalized, (int)base->d16,
                                   (int)base->kp, "gq", base->center);
      bf_search(*base, Q, (int)k, metric, 0, out_ids, out_dist, nullptr);
    };
    // an eager run first: it sets the kernels' attributes outside the capture
    // (and computes this batch's answer), then the same search is captured
    run();
    MANN_HIP(hipStreamSynchronize(c.stream));
    g->ci.fallback_slots = F;
    c.capture = &g->ci;
    MANN_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
      run();
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(c.stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    MANN_HIP(hipStreamEndCapture(c.stream, &g->graph));
    MANN_REQUIRE(g->ci.filled, ERR_ASSERTION, "the captured search recorded no fallback");
    MANN_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

int mann_graph_launch(mann_graph_t g) {
  return guarded([&] {
    MANN_REQUIRE(g != nullptr, ERR_NULL, "null graph");
    MANN_HIP(hipSetDevice(g->device));
    ThreadCtx& c = ctx();
    hipStream_t st = c.stream;
    MANN_HIP(hipGraphLaunch(g->exec, st));
    int hf[8], hs[kStatSlots * 4];
    MANN_HIP(hipMemcpyAsync(hf, g->ci.flags, sizeof(hf), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipMemcpyAsync(hs, g->ci.stats, sizeof(hs), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    // flagged queries beyond the graph's fallback slots: the rest of the
    // exact fallback, eagerly, with the graph's own partial buffers
    const int F = g->ci.fallback_slots;
    if (hf[1] > F) {
      ExactArgs e = g->ci.e;
      for (int b = F; b < hf[1]; b += F) {
        e.slot_base = b;
        launch_exact_topk(e, std::min(F, hf[1] - b), st);
      }
      MANN_HIP(hipStreamSynchronize(st));
    }
    MANN_HIP(hipGetLastError());
    SearchStats& S = last_stats();
    S = SearchStats();
    S.pass2 = hf[0];
    S.exact = hf[1];
    add_stats(hs, S);
  });
}

int mann_graph_destroy(mann_graph_t g) {
  return guarded([&] { delete g; });
}

int mann_ivf_search_graph_create(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq,
                                 int64_t ldq, int64_t k, int64_t nprobe, int metric,
                                 const float* route_dist, int64_t* out_ids, float* out_dist,
                                 mann_graph_t* out) {
  return guarded([&] {
    MANN_REQUIRE(out != nullptr, ERR_NULL, "null output handle");
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    check_metric(metric);
    MANN_REQUIRE((metric == COSINE) == (ivf->normalized != 0), ERR_ARGUMENT,
                 "index was built for another metric");
    MANN_REQUIRE(!ivf->flat, ERR_ARGUMENT,
                 "a captured IVF search merges in the batched SimpleMerge order only (not the "
                 "IVF-Flat heap)");
    MANN_REQUIRE(d == ivf->d, ERR_DIMENSION,
                 "Expected queries with " + std::to_string(ivf->d) + " rows, got " +
                     std::to_string(d));
    MANN_REQUIRE(nq > 0 && k > 0, ERR_ARGUMENT, "a captured search needs queries and k > 0");
    MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT, "nprobe must be positive (TopKRouting: k must be positive)");
    MANN_REQUIRE(queries && out_ids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(ldq >= d, ERR_ARGUMENT, "leading dimension smaller than d");
    MANN_REQUIRE(is_device_ptr(queries) && is_device_ptr(out_ids) && is_device_ptr(out_dist) &&
                     (route_dist == nullptr || is_device_ptr(route_dist)),
                 ERR_ARGUMENT, "a captured search takes device queries, routes and outputs only");
    ThreadCtx& c = ctx();
    auto g = std::make_unique<mann_graph>();
    g->device = c.device;
    g->nq = (int)nq;
    g->ci.fallback_slots = (int)std::min<int64_t>(nq, 64);
    struct Restore {
      ThreadCtx& c;
      ~Restore() {
        c.owned = nullptr;
        c.capture = nullptr;
      }
    } restore{c};
    c.owned = &g->owned;
    c.capture = &g->ci;
    auto run = [&] {
      QueryOps Q = prepare_queries(queries, d, nq, ldq, ivf->normalized, (int)ivf->d16,
                                   (int)ivf->kp, "ivq");
      ivf_search(*ivf, Q, (int)k, (int)nprobe, metric, route_dist, out_ids, out_dist, 0.0f, 0,
                 0);
    };
    // an eager run of the device-gated search sizes the graph's buffers and
    // sets the kernels' attributes; the capture then allocates nothing
    run();
    MANN_HIP(hipStreamSynchronize(c.stream));
    g->ci.filled = false;
    MANN_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
      run();
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(c.stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    MANN_HIP(hipStreamEndCapture(c.stream, &g->graph));
    MANN_REQUIRE(g->ci.filled, ERR_ASSERTION, "the captured search recorded no fallback");
    MANN_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

// ---- autotuner ----------------------------------------------------------------------

int mann_autotune_bf_knn(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                         int64_t ldq, int64_t k, int metric, int64_t* out_choice,
                         double* out_ms) {
  return guarded([&] {
    MANN_REQUIRE(base != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    MANN_REQUIRE((metric == COSINE) == (base->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    MANN_HIP(hipSetDevice(base->device));
    ThreadCtx& c = ctx();
    QueryOps Q;
    int exclude_self = 0;
    if (queries == nullptr) {  // the kNN graph's self-join
      Q = dataset_as_queries(*base);
      exclude_self = 1;
      nq = base->n;
    } else {
      MANN_REQUIRE(d == base->d, ERR_DIMENSION,
                   "Expected queries with " + std::to_string(base->d) + " rows, got " +
                       std::to_string(d));
      MANN_REQUIRE(nq > 0 && ldq >= d, ERR_ARGUMENT, "bad query matrix");
      Q = prepare_queries(queries, d, nq, ldq, base->normalized, (int)base->d16,
                          (int)base->kp, "atq", base->center);
    }
    MANN_REQUIRE(k > 0 && k + exclude_self <= select_exmax() / 2, ERR_ARGUMENT,
                 "k out of range for the GPU search");
    int64_t* ids = c.buf("at_ids").as<int64_t>((size_t)nq * k);
    float* dist = c.buf("at_dist").as<float>((size_t)nq * k);
    AutotuneResult R = autotune_bf(*base, Q, (int)k, metric, exclude_self, ids, dist);
    if (out_choice) {
      out_choice[0] = R.best.scan_cfg;
      out_choice[1] = R.best.emit_per_k;
      out_choice[2] = R.best.shortlist_mult;
      out_choice[3] = R.best.scan8_cfg;
      out_choice[4] = R.best.fp8_emit_per_k;
      out_choice[5] = R.trials;
    }
    if (out_ms) *out_ms = R.ms;
  });
}

static QueryOps ivf_queries(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq,
                            int64_t ldq);

int mann_autotune_ivf(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                      int64_t k, int64_t nprobe, int metric, const float* route_dist,
                      int64_t* out_choice, double* out_ms) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    check_metric(metric);
    MANN_REQUIRE((metric == COSINE) == (ivf->normalized != 0), ERR_ARGUMENT,
                 "index was built for another metric");
    MANN_REQUIRE(!ivf->flat, ERR_ARGUMENT, "the IVF autotuner covers MultiLevelIndex searches");
    MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT, "nprobe must be positive (TopKRouting: k must be positive)");
    MANN_REQUIRE(k > 0 && k <= select_exmax() / 2, ERR_ARGUMENT,
                 "k must lie in [1, " + std::to_string(select_exmax() / 2) + "]");
    QueryOps Q = ivf_queries(ivf, queries, d, nq, ldq);
    MANN_REQUIRE(nq > 0, ERR_ARGUMENT, "autotuning needs queries");
    ThreadCtx& c = ctx();
    const float* R = route_dist;
    if (R && !is_device_ptr(R)) {
      float* rb = c.buf("at_R").as<float>((size_t)nq * ivf->nlist);
      MANN_HIP(hipMemcpyAsync(rb, route_dist, sizeof(float) * nq * ivf->nlist,
                              hipMemcpyHostToDevice, c.stream));
      R = rb;
    }
    int64_t* ids = c.buf("at_ids").as<int64_t>((size_t)nq * k);
    float* dist = c.buf("at_dist").as<float>((size_t)nq * k);
    AutotuneResult A = autotune_ivf(*ivf, Q, (int)k, (int)nprobe, metric, R, ids, dist);
    if (out_choice) {
      out_choice[0] = A.best.ivf_cfg;
      out_choice[1] = A.best.ivf_rows_item;
      out_choice[2] = A.best.ivf_tau_cfg;
      out_choice[3] = A.best.shortlist_mult;
      out_choice[4] = A.trials;
    }
    if (out_ms) *out_ms = A.ms;
    timer_flush();
  });
}

int mann_autotune_clear(void) {
  return guarded([&] { shape_tuning_clear(); });
}

int mann_knn_graph(mann_dataset_t data, int64_t k, int include_self, int metric,
                   int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(data != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    const int64_t n = data->n;
    const int64_t maxnb = n - (include_self ? 0 : 1);
    MANN_REQUIRE(k > 0 && k <= maxnb, ERR_ARGUMENT,
                 "k must satisfy 0 < k <= " + std::to_string(maxnb) + " for dataset with " +
                     std::to_string(n) + " points");
    MANN_REQUIRE(out_ids != nullptr, ERR_NULL, "null pointer");
    MANN_REQUIRE((metric == COSINE) == (data->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    QueryOps Q = dataset_as_queries(*data);
    OutBuf<int64_t> oi(out_ids, (size_t)(n * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(n * k), "out_dist");
    int* cnt = ctx().buf("graph_cnt").as<int>(n);
    bf_search(*data, Q, (int)k, metric, include_self ? 0 : 1, oi.dev, od.dev, cnt);
    // _prepare_neighbors throws when a row comes back short (knn_graph.jl:122-128)
    std::vector<int> hc(n);
    MANN_HIP(hipMemcpy(hc.data(), cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i)
      MANN_REQUIRE(hc[i] == k, ERR_ARGUMENT,
                   "Index could only supply " + std::to_string(hc[i]) +
                       " usable neighbors for point " + std::to_string(i + 1) +
                       ". Consider lowering k or setting include_self=true.");
    oi.finish();
    if (out_dist) od.finish();
    timer_flush();
  });
}

int mann_knn_candidates(mann_dataset_t data, int64_t candidate_k, int metric, int64_t* out_ids,
                        float* out_dist, int64_t* out_count) {
  return guarded([&] {
    MANN_REQUIRE(data != nullptr, ERR_NULL, "null dataset");
    check_metric(metric);
    MANN_REQUIRE(candidate_k > 0, ERR_ARGUMENT, "candidate_k must be positive");
    MANN_REQUIRE(out_ids != nullptr, ERR_NULL, "null pointer");
    MANN_REQUIRE((metric == COSINE) == (data->normalized != 0), ERR_ARGUMENT,
                 "dataset was prepared for another metric");
    const int64_t n = data->n, k = candidate_k;
    QueryOps Q = dataset_as_queries(*data);
    OutBuf<int64_t> oi(out_ids, (size_t)(n * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(n * k), "out_dist");
    int* cnt = ctx().buf("cand_cnt").as<int>(n);
    // query k + 1 (at most n results), drop the point itself by id, keep the
    // first k: a short row is not an error here
    bf_search(*data, Q, (int)k, metric, 1, oi.dev, od.dev, cnt);
    if (out_count) {
      std::vector<int> hc(n);
      MANN_HIP(hipMemcpy(hc.data(), cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
      std::vector<int64_t> h64(hc.begin(), hc.end());
      MANN_HIP(hipMemcpy(out_count, h64.data(), sizeof(int64_t) * n, hipMemcpyDefault));
    }
    oi.finish();
    if (out_dist) od.finish();
    timer_flush();
  });
}

int mann_knn_graph_symmetrize(const int64_t* adj, int64_t n, int64_t k, int64_t* out_offsets,
                              int64_t* out_ids, int64_t* out_total, int64_t* out_max_degree) {
  return guarded([&] {
    MANN_REQUIRE(adj && out_offsets && out_ids, ERR_NULL, "null pointer");
    MANN_REQUIRE(n >= 0 && k >= 0, ERR_ARGUMENT, "negative size");
    const bool dev = is_device_ptr(adj);
    MANN_REQUIRE(dev == is_device_ptr(out_offsets) && dev == is_device_ptr(out_ids),
                 ERR_ARGUMENT, "adjacency and outputs must all be host or all be device memory");
    if (dev) {
      knn_graph_symmetrize_gpu(adj, n, k, out_offsets, out_ids, out_total, out_max_degree);
      timer_flush();
      return;
    }
    // host arrays: stage through device buffers
    ThreadCtx& c = ctx();
    hipStream_t st = c.stream;
    const int64_t m = 2 * n * k;
    int64_t* da = c.buf("sym_adj").as<int64_t>(std::max<int64_t>(n * k, 1));
    int64_t* doff = c.buf("sym_off").as<int64_t>(n + 1);
    int64_t* dids = c.buf("sym_ids").as<int64_t>(std::max<int64_t>(m, 1));
    if (n * k > 0)
      MANN_HIP(hipMemcpyAsync(da, adj, sizeof(int64_t) * n * k, hipMemcpyHostToDevice, st));
    int64_t tot = 0, maxd = 0;
    knn_graph_symmetrize_gpu(da, n, k, doff, dids, &tot, &maxd);
    MANN_HIP(hipMemcpyAsync(out_offsets, doff, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost,
                            st));
    if (tot > 0)
      MANN_HIP(hipMemcpyAsync(out_ids, dids, sizeof(int64_t) * tot, hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    if (out_total) *out_total = tot;
    if (out_max_degree) *out_max_degree = maxd;
    timer_flush();
  });
}

// Upload (if needed) a k x d centroid matrix into a k x d16 device buffer.
static float* stage_centroids(const float* C, int64_t k, const mann_dataset& ds, const char* tag) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  float* dst = c.buf(tag).as<float>((size_t)k * ds.d16);
  const float* src = C;
  if (!is_device_ptr(C)) {
    float* stg = c.buf(std::string(tag) + "_stage").as<float>((size_t)k * ds.d);
    MANN_HIP(hipMemcpyAsync(stg, C, sizeof(float) * k * ds.d, hipMemcpyHostToDevice, st));
    src = stg;
  }
  copy_rows(src, ds.d, dst, ds.d16, (int)ds.d, k, st);
  return dst;
}

// Write k x d16 device centroids into the caller's k x d matrix.
static void emit_centroids(const float* C, int64_t k, const mann_dataset& ds, float* out) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  OutBuf<float> ob(out, (size_t)(k * ds.d), "out_centroids");
  copy_rows(C, ds.d16, ob.dev, ds.d, (int)ds.d, k, st);
  ob.finish();
}

int mann_kmeans_assign(mann_dataset_t X, const float* centroids, int64_t k, int64_t* out_assign,
                       float* out_mindist) {
  return guarded([&] {
    MANN_REQUIRE(X && centroids && out_assign, ERR_NULL, "null pointer");
    MANN_REQUIRE(k > 0, ERR_ASSERTION, "Number of clusters must be positive");
    MANN_REQUIRE(!X->normalized, ERR_ARGUMENT, "k-means runs on a Euclidean dataset");
    hipStream_t st = ctx().stream;
    const int64_t n = X->n;
    float* C = stage_centroids(centroids, k, *X, "ka_c");
    int* assign = ctx().buf("ka_assign").as<int>(n);
    float* md = ctx().buf("ka_md").as<float>(n);
    long long ovf = 0;
    kmeans_estep(*X, C, (int)k, assign, md, false, &ovf);
    last_stats() = SearchStats();
    last_stats().overflow = ovf;
    OutBuf<int64_t> oa(out_assign, (size_t)n, "out_assign");
    ids_to_i64(assign, n, oa.dev, st);
    oa.finish();
    if (out_mindist) {
      OutBuf<float> om(out_mindist, (size_t)n, "out_mindist");
      MANN_HIP(hipMemcpyAsync(om.dev, md, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
      om.finish();
    }
    MANN_HIP(hipStreamSynchronize(st));
    timer_flush();
  });
}

int mann_kmeanspp(mann_dataset_t X, int64_t k, uint64_t seed, float* out_centroids,
                  int64_t* out_idx) {
  return guarded([&] {
    MANN_REQUIRE(X && out_centroids, ERR_NULL, "null pointer");
    MANN_REQUIRE(k > 0, ERR_ASSERTION, "Number of clusters must be positive");
    MANN_REQUIRE(k <= X->n, ERR_ASSERTION,
                 "Cannot initialize " + std::to_string(k) + " clusters with only " +
                     std::to_string(X->n) + " points");
    hipStream_t st = ctx().stream;
    float* C = ctx().buf("pp_c").as<float>((size_t)k * X->d16);
    int* picked = ctx().buf("pp_picked").as<int>(k);
    kmeans_pp(*X, (int)k, seed, C, picked);
    emit_centroids(C, k, *X, out_centroids);
    if (out_idx) {
      OutBuf<int64_t> oi(out_idx, (size_t)k, "out_idx");
      ids_to_i64(picked, k, oi.dev, st);
      oi.finish();
    }
    MANN_HIP(hipStreamSynchronize(st));
    timer_flush();
  });
}

int mann_kmeans_fit(mann_dataset_t X, int64_t k, int64_t max_iters, double tol, int init,
                    uint64_t seed, const float* init_centroids, int64_t subsample,
                    float* out_centroids, int64_t* out_assign, int64_t* out_iters) {
  return guarded([&] {
    MANN_REQUIRE(X && out_centroids, ERR_NULL, "null pointer");
    // KMeansTransform constructor assertions (KMeansTransform.jl:66-69) and fit! (:110)
    MANN_REQUIRE(k > 0, ERR_ASSERTION, "Number of clusters must be positive");
    MANN_REQUIRE(max_iters > 0, ERR_ASSERTION, "max_iters must be positive");
    MANN_REQUIRE(tol > 0, ERR_ASSERTION, "tol must be positive");
    MANN_REQUIRE(k <= X->n, ERR_ASSERTION,
                 "Cannot fit " + std::to_string(k) + " clusters with only " +
                     std::to_string(X->n) + " points");
    MANN_REQUIRE(init >= 0 && init <= 2, ERR_ARGUMENT,
                 "Unknown initialization strategy");
    MANN_REQUIRE(!X->normalized, ERR_ARGUMENT, "k-means runs on a Euclidean dataset");
    MANN_REQUIRE(subsample >= 0, ERR_ARGUMENT, "subsample_size must be positive or nothing");
    ThreadCtx& c = ctx();
    hipStream_t st = c.stream;
    const int64_t n = X->n;
    // fit on a uniform subsample of max(subsample, k) distinct points in the
    // order drawn (X[:, randperm(rng, n)[1:sample_n]], KMeansTransform.jl:113-120),
    // then assign every point with one E-step (:137-146)
    const bool use_sub = subsample > 0 && subsample < n;
    std::unique_ptr<mann_dataset> sub;
    const mann_dataset* T = X;
    uint64_t fit_seed = seed;
    if (use_sub) {
      const int64_t m = std::max<int64_t>(subsample, k);
      std::vector<int> idx = random_indices(n, (int)m, seed);
      int* didx = c.buf("kf_sidx").as<int>(m);
      float* rows = c.buf("kf_srows").as<float>((size_t)m * X->d16);
      MANN_HIP(hipMemcpyAsync(didx, idx.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
      gather_rows_f32(X->x32, didx, (int)m, (int)X->d16, rows, st);
      sub.reset(create_dataset(rows, X->d, m, X->d16, EUCLIDEAN));
      T = sub.get();
      fit_seed = seed * 0x9E3779B97F4A7C15ull + 1;  // an independent stream for the fit
    }
    const int64_t nt = T->n;
    float* C = c.buf("kf_c").as<float>((size_t)k * X->d16);
    if (init == 0) {
      MANN_REQUIRE(init_centroids != nullptr, ERR_NULL, "null initial centroids");
      float* src = stage_centroids(init_centroids, k, *X, "kf_init");
      MANN_HIP(hipMemcpyAsync(C, src, sizeof(float) * k * X->d16, hipMemcpyDeviceToDevice, st));
    } else if (init == 1) {
      int* picked = c.buf("pp_picked").as<int>(k);
      kmeans_pp(*T, (int)k, fit_seed, C, picked);
    } else {
      std::vector<int> idx = random_indices(nt, (int)k, fit_seed);
      int* didx = c.buf("kf_ridx").as<int>(k);
      MANN_HIP(hipMemcpyAsync(didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
      gather_rows_f32(T->x32, didx, (int)k, (int)X->d16, C, st);
    }
    int* assign = c.buf("kf_assign").as<int>(n);
    float* md = c.buf("kf_md").as<float>(n);
    long long ovf = 0;
    int64_t iters = kmeans_lloyd(*T, C, (int)k, max_iters, tol, assign, md, &ovf);
    if (use_sub) {
      long long ovf2 = 0;
      kmeans_estep(*X, C, (int)k, assign, md, false, &ovf2);
      ovf += ovf2;
    }
    last_stats() = SearchStats();
    last_stats().overflow = ovf;
    emit_centroids(C, k, *X, out_centroids);
    if (out_assign) {
      OutBuf<int64_t> oa(out_assign, (size_t)n, "out_assign");
      ids_to_i64(assign, n, oa.dev, st);
      oa.finish();
    }
    if (out_iters) *out_iters = iters;
    MANN_HIP(hipStreamSynchronize(st));
    timer_flush();
  });
}

int mann_ivf_build(mann_dataset_t X, const float* centroids, int64_t nlist,
                   const int64_t* assign, mann_ivf_t* out) {
  return guarded([&] {
    MANN_REQUIRE(X && centroids && assign && out, ERR_NULL, "null pointer");
    MANN_REQUIRE(nlist > 0, ERR_ASSERTION, "Number of clusters must be positive");
    std::vector<int64_t> ha(X->n);
    if (is_device_ptr(assign)) {
      MANN_HIP(hipMemcpy(ha.data(), assign, sizeof(int64_t) * X->n, hipMemcpyDeviceToHost));
    } else {
      memcpy(ha.data(), assign, sizeof(int64_t) * X->n);
    }
    float* C = stage_centroids(centroids, nlist, *X, "ivb_c");
    *out = ivf_build(*X, C, (int)nlist, ha.data());
  });
}

int mann_ivf_flat_build(mann_dataset_t X, const float* centroids, int64_t nlist,
                        const int64_t* assign, int centroid_metric, mann_ivf_t* out) {
  return guarded([&] {
    MANN_REQUIRE(X && centroids && assign && out, ERR_NULL, "null pointer");
    MANN_REQUIRE(nlist > 0, ERR_ARGUMENT, "nlist must be positive");
    check_metric(centroid_metric);
    std::vector<int64_t> ha(X->n);
    if (is_device_ptr(assign)) {
      MANN_HIP(hipMemcpy(ha.data(), assign, sizeof(int64_t) * X->n, hipMemcpyDeviceToHost));
    } else {
      memcpy(ha.data(), assign, sizeof(int64_t) * X->n);
    }
    for (int64_t j = 0; j < X->n; ++j)
      MANN_REQUIRE(ha[j] >= 1 && ha[j] <= nlist, ERR_ARGUMENT,
                   "Assignment out of bounds: " + std::to_string(ha[j]));
    float* C = stage_centroids(centroids, nlist, *X, "ivb_c");
    *out = ivf_build(*X, C, (int)nlist, ha.data(), true, centroid_metric);
  });
}

int mann_ivf_destroy(mann_ivf_t ivf) {
  return guarded([&] { delete ivf; });
}

int mann_ivf_info(mann_ivf_t ivf, int64_t* n_buckets, int64_t* list_sizes) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    if (n_buckets) *n_buckets = ivf->nb;
    if (list_sizes)
      for (int b = 0; b < ivf->nb; ++b) list_sizes[b] = ivf->h_off[b + 1] - ivf->h_off[b];
  });
}

int mann_ivf_list_ids(mann_ivf_t ivf, int64_t* out_ids) {
  return guarded([&] {
    MANN_REQUIRE(ivf && out_ids, ERR_NULL, "null pointer");
    std::vector<int> p(ivf->n);
    MANN_HIP(hipMemcpy(p.data(), ivf->perm, sizeof(int) * ivf->n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ivf->n; ++i) out_ids[i] = (int64_t)p[i] + 1;
  });
}

static QueryOps ivf_queries(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq,
                            int64_t ldq) {
  MANN_REQUIRE(d == ivf->d, ERR_DIMENSION,
               "Expected queries with " + std::to_string(ivf->d) + " rows, got " +
                   std::to_string(d));
  MANN_REQUIRE(nq >= 0, ERR_ARGUMENT, "negative query count");
  MANN_REQUIRE(nq == 0 || (queries && ldq >= d), ERR_ARGUMENT, "bad query matrix");
  return prepare_queries(queries, d, nq, ldq, ivf->normalized, (int)ivf->d16, (int)ivf->kp,
                         "ivq");
}

int mann_ivf_route(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                   float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(ivf && out_dist, ERR_NULL, "null pointer");
    QueryOps Q = ivf_queries(ivf, queries, d, nq, ldq);
    if (nq == 0) return;
    OutBuf<float> od(out_dist, (size_t)(nq * ivf->nlist), "out_route");
    ivf_route(*ivf, Q, od.dev);
    od.finish();
    MANN_HIP(hipStreamSynchronize(ctx().stream));
    timer_flush();
  });
}

static void probe_out(mann_ivf_t ivf, const float* route_dist, int64_t nq, int64_t nprobe,
                      float mass, int64_t* out_buckets) {
  MANN_REQUIRE(ivf && route_dist && out_buckets, ERR_NULL, "null pointer");
  MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT, "nprobe must be positive (TopKRouting: k must be positive)");
  if (nq == 0) return;
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const float* R = route_dist;
  if (!is_device_ptr(route_dist)) {
    float* rb = c.buf("pr_R").as<float>((size_t)nq * ivf->nlist);
    MANN_HIP(hipMemcpyAsync(rb, route_dist, sizeof(float) * nq * ivf->nlist,
                            hipMemcpyHostToDevice, st));
    R = rb;
  }
  int* probes = c.buf("pr_probes").as<int>((size_t)nq * nprobe);
  int* pcount = c.buf("pr_pcount").as<int>(nq);
  int* prows = c.buf("pr_prows").as<int>(nq);
  ivf_probe(*ivf, R, (int)nq, (int)nprobe, probes, pcount, prows, mass);
  std::vector<int> hp((size_t)nq * nprobe);
  MANN_HIP(hipMemcpyAsync(hp.data(), probes, sizeof(int) * hp.size(), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> tmp(hp.size());
  for (size_t i = 0; i < hp.size(); ++i) tmp[i] = hp[i] + 1;  // -1 -> 0 (empty slot)
  if (is_device_ptr(out_buckets))
    MANN_HIP(hipMemcpy(out_buckets, tmp.data(), sizeof(int64_t) * tmp.size(),
                       hipMemcpyHostToDevice));
  else
    memcpy(out_buckets, tmp.data(), sizeof(int64_t) * tmp.size());
}

int mann_ivf_probe(mann_ivf_t ivf, const float* route_dist, int64_t nq, int64_t nprobe,
                   int64_t* out_buckets) {
  return guarded([&] { probe_out(ivf, route_dist, nq, nprobe, 0.0f, out_buckets); });
}

static void search_ex(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                      int64_t k, int64_t nprobe, int metric, const float* route_dist, int merge,
                      int probe_order, int64_t* out_ids, float* out_dist);

int mann_ivf_search(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                    int64_t k, int64_t nprobe, int metric, const float* route_dist,
                    int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    search_ex(ivf, queries, d, nq, ldq, k, nprobe, metric, route_dist, MANN_MERGE_SIMPLE, 0,
              out_ids, out_dist);
  });
}

int mann_ivf_search_ex(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                       int64_t k, int64_t nprobe, int metric, const float* route_dist, int merge,
                       int probe_order, int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(merge == MANN_MERGE_SIMPLE || merge == MANN_MERGE_DISJOINT, ERR_ARGUMENT,
                 "unknown merge strategy");
    search_ex(ivf, queries, d, nq, ldq, k, nprobe, metric, route_dist, merge, probe_order != 0,
              out_ids, out_dist);
  });
}

static void search_ex(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                      int64_t k, int64_t nprobe, int metric, const float* route_dist, int merge,
                      int probe_order, int64_t* out_ids, float* out_dist) {
  {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    check_metric(metric);
    MANN_REQUIRE((metric == COSINE) == (ivf->normalized != 0), ERR_ARGUMENT,
                 "index was built for another metric");
    // IVF-Flat clamps nprobe to [1, nlist] (ivf_flat.jl:146); TopKRouting
    // rejects k <= 0 (routing.jl:39-42)
    if (ivf->flat) nprobe = std::max<int64_t>(nprobe, 1);
    MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT, "nprobe must be positive (TopKRouting: k must be positive)");
    QueryOps Q = ivf_queries(ivf, queries, d, nq, ldq);
    if (nq == 0 || k <= 0) {
      last_stats() = SearchStats();
      return;
    }
    MANN_REQUIRE(out_ids && out_dist, ERR_NULL, "null pointer");
    ThreadCtx& c = ctx();
    const float* R = route_dist;
    if (R && !is_device_ptr(R)) {
      float* rb = c.buf("sr_R").as<float>((size_t)nq * ivf->nlist);
      MANN_HIP(hipMemcpyAsync(rb, route_dist, sizeof(float) * nq * ivf->nlist,
                              hipMemcpyHostToDevice, c.stream));
      R = rb;
    }
    OutBuf<int64_t> oi(out_ids, (size_t)(nq * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(nq * k), "out_dist");
    ivf_search(*ivf, Q, (int)k, (int)nprobe, metric, R, oi.dev, od.dev, 0.0f, merge,
               probe_order);
    oi.finish();
    od.finish();
    timer_flush();
  }
}

int mann_ivf_hnsw_build(mann_ivf_t ivf, int64_t M, int64_t ef_construction, int neighbor_policy,
                        uint64_t seed) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    MANN_REQUIRE(M > 0 && M < (1 << 30), ERR_ARGUMENT, "M must be positive");
    MANN_REQUIRE(ef_construction >= 1 && ef_construction < (1 << 30), ERR_ARGUMENT,
                 "ef_construction must be at least 1");
    ivf_hnsw_build(*ivf, (int)M, (int)ef_construction, neighbor_policy, seed);
  });
}

int mann_ivf_hnsw_search(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                         int64_t k, int64_t nprobe, int64_t ef_search, const float* route_dist,
                         int merge, int probe_order, int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    MANN_REQUIRE(merge == MANN_MERGE_SIMPLE || merge == MANN_MERGE_DISJOINT, ERR_ARGUMENT,
                 "unknown merge strategy");
    MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT,
                 "nprobe must be positive (TopKRouting: k must be positive)");
    MANN_REQUIRE(ef_search >= 1 && ef_search < (1 << 30), ERR_ARGUMENT,
                 "ef_search must be at least 1");
    QueryOps Q = ivf_queries(ivf, queries, d, nq, ldq);
    if (nq == 0 || k <= 0) return;
    MANN_REQUIRE(out_ids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(k < (1 << 30), ERR_ARGUMENT, "k too large");
    ThreadCtx& c = ctx();
    const float* R = route_dist;
    if (R && !is_device_ptr(R)) {
      float* rb = c.buf("hn_Rin").as<float>((size_t)nq * ivf->nlist);
      MANN_HIP(hipMemcpyAsync(rb, route_dist, sizeof(float) * nq * ivf->nlist,
                              hipMemcpyHostToDevice, c.stream));
      R = rb;
    }
    OutBuf<int64_t> oi(out_ids, (size_t)(nq * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(nq * k), "out_dist");
    ivf_hnsw_search(*ivf, Q, (int)k, (int)std::min<int64_t>(nprobe, 1 << 30),
                    (int)ef_search, R, merge, probe_order != 0, oi.dev, od.dev);
    oi.finish();
    od.finish();
    MANN_HIP(hipStreamSynchronize(c.stream));
    timer_flush();
  });
}

int mann_ivf_hnsw_set_graph(mann_ivf_t ivf, int64_t layers, int64_t M, const int32_t* nbr,
                            const int32_t* deg, const int32_t* entry, const int32_t* max_layer,
                            const int32_t* level) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    MANN_REQUIRE(nbr && deg && entry && max_layer, ERR_NULL, "null pointer");
    MANN_REQUIRE(layers >= 1 && layers <= 64 && M >= 1 && M <= HN_M_MAX, ERR_ARGUMENT,
                 "layers must lie in [1, 64] and M in [1, " + std::to_string(HN_M_MAX) + "]");
    for (const void* p : {(const void*)nbr, (const void*)deg, (const void*)entry,
                          (const void*)max_layer, (const void*)level})
      MANN_REQUIRE(!p || !is_device_ptr(p), ERR_ARGUMENT, "the graph arrays are host arrays");
    ivf_hnsw_set_graph(*ivf, (int)layers, (int)M, nbr, deg, entry, max_layer, level);
  });
}

int mann_ivf_hnsw_graph(mann_ivf_t ivf, int64_t* out_layers, int64_t* out_M, int32_t* out_nbr,
                        int32_t* out_deg, int32_t* out_entry, int32_t* out_max_layer,
                        int32_t* out_level) {
  return guarded([&] {
    MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
    MANN_REQUIRE(ivf->hn_nbr0 != nullptr, ERR_ARGUMENT, "the index has no HNSW terminals");
    const int64_t n = ivf->n, L = ivf->hn_layers, cap = ivf->hn_M + 1, n1 = ivf->hn_n1;
    if (out_layers) *out_layers = L;
    if (out_M) *out_M = ivf->hn_M;
    MANN_HIP(hipStreamSynchronize(ctx().stream));
    auto get = [&](const int* src, int64_t count) {
      std::vector<int> h((size_t)std::max<int64_t>(count, 1));
      if (count > 0) MANN_HIP(hipMemcpy(h.data(), src, sizeof(int) * count, hipMemcpyDeviceToHost));
      return h;
    };
    if (out_entry || out_max_layer) {
      std::vector<int> e = get(ivf->hn_entry, ivf->nb), m = get(ivf->hn_maxl, ivf->nb);
      for (int b = 0; b < ivf->nb; ++b) {
        if (out_entry) out_entry[b] = e[b];
        if (out_max_layer) out_max_layer[b] = m[b];
      }
    }
    if (out_level) {
      std::vector<int> lv = get(ivf->hn_level, n);
      for (int64_t i = 0; i < n; ++i) out_level[i] = lv[i];
    }
    if (!out_nbr && !out_deg) return;
    // dense [layer][row][M + 1] / [layer][row], rows absent from a layer zero
    std::vector<int> nb0 = get(ivf->hn_nbr0, n * cap), dg0 = get(ivf->hn_deg0, n);
    std::vector<int> nbu = get(ivf->hn_nbru, (L - 1) * n1 * cap), dgu = get(ivf->hn_degu, (L - 1) * n1);
    std::vector<int> up = get(ivf->hn_upslot, n);
    for (int64_t l = 0; l < L; ++l)
      for (int64_t r = 0; r < n; ++r) {
        int dg = 0;
        const int* src = nullptr;
        if (l == 0) {
          dg = dg0[r];
          src = nb0.data() + r * cap;
        } else if (up[r] >= 0) {
          const int64_t s = (l - 1) * n1 + up[r];
          dg = dgu[s];
          src = nbu.data() + s * cap;
        }
        if (out_deg) out_deg[l * n + r] = dg;
        if (out_nbr)
          for (int64_t j = 0; j < cap; ++j) out_nbr[(l * n + r) * cap + j] = src ? src[j] : 0;
      }
  });
}

int mann_route_distances(const float* centroids, int64_t nlist, int64_t d, const float* queries,
                         int64_t nq, int64_t ldq, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(centroids && out_dist, ERR_NULL, "null pointer");
    MANN_REQUIRE(nlist > 0 && d > 0, ERR_ARGUMENT, "empty centroid matrix");
    if (nq == 0) return;
    MANN_REQUIRE(queries && ldq >= d, ERR_ARGUMENT, "bad query matrix");
    ThreadCtx& c = ctx();
    hipStream_t st = c.stream;
    const int64_t d16 = round_up(d, 16), kp = round_up(d, 32);
    QueryOps Q = prepare_queries(queries, d, nq, ldq, 0, (int)d16, (int)kp, "rdq");
    float* C = c.buf("rd_c").as<float>((size_t)nlist * d16);
    const float* src = centroids;
    if (!is_device_ptr(centroids)) {
      float* stg = c.buf("rd_c_stage").as<float>((size_t)nlist * d);
      MANN_HIP(hipMemcpyAsync(stg, centroids, sizeof(float) * nlist * d, hipMemcpyHostToDevice,
                              st));
      src = stg;
    }
    copy_rows(src, d, C, d16, (int)d, nlist, st);
    float* cn2 = c.buf("rd_cn2").as<float>(nlist);
    launch_a16_norms(C, nlist, (int)d16, cn2, st);
    OutBuf<float> od(out_dist, (size_t)(nq * nlist), "out_route");
    launch_route_dist(Q.q32, Q.qn2, (int)nq, C, cn2, (int)nlist, (int)d16, od.dev, st);
    od.finish();
    MANN_HIP(hipStreamSynchronize(st));
  });
}

// Dense device copy (row stride d) of an n x d row matrix with stride ld.
static const float* stage_rows(const float* X, int64_t d, int64_t n, int64_t ld, const char* tag,
                               int64_t* ld_out) {
  ThreadCtx& c = ctx();
  const float* src = X;
  if (!is_device_ptr(X)) {
    float* stg = c.buf(std::string(tag) + "_stage").as<float>((size_t)(ld * (n - 1) + d));
    MANN_HIP(hipMemcpyAsync(stg, X, sizeof(float) * (ld * (n - 1) + d), hipMemcpyHostToDevice,
                            c.stream));
    src = stg;
  }
  float* dst = c.buf(tag).as<float>((size_t)(n * d));
  copy_rows(src, ld, dst, d, (int)d, n, c.stream);
  *ld_out = d;
  return dst;
}

static const int* stage_labels(const int64_t* labels, int64_t count, const char* tag) {
  ThreadCtx& c = ctx();
  const int64_t* src = labels;
  if (!is_device_ptr(labels)) {
    int64_t* stg = c.buf(std::string(tag) + "_stage").as<int64_t>(count);
    MANN_HIP(hipMemcpyAsync(stg, labels, sizeof(int64_t) * count, hipMemcpyHostToDevice,
                            c.stream));
    src = stg;
  }
  int* dst = c.buf(tag).as<int>(count);
  labels_to_i32(src, count, dst, c.stream);
  return dst;
}

int mann_router_create(int64_t d, int64_t hidden, int64_t nlist, uint64_t seed,
                       mann_router_t* out) {
  return guarded([&] {
    MANN_REQUIRE(out != nullptr, ERR_NULL, "null output handle");
    MANN_REQUIRE(d > 0 && hidden > 0 && nlist > 0, ERR_ARGUMENT, "router sizes must be positive");
    *out = router_create((int)d, (int)hidden, (int)nlist, seed);
  });
}

int mann_router_destroy(mann_router_t r) {
  return guarded([&] { delete r; });
}

int mann_router_params(mann_router_t r, float* out, int64_t* count) {
  return guarded([&] {
    MANN_REQUIRE(r != nullptr, ERR_NULL, "null router");
    if (count) *count = r->nparam;
    if (out) MANN_HIP(hipMemcpy(out, r->w, sizeof(float) * r->nparam, hipMemcpyDefault));
  });
}

int mann_router_set_params(mann_router_t r, const float* params) {
  return guarded([&] {
    MANN_REQUIRE(r && params, ERR_NULL, "null pointer");
    MANN_HIP(hipMemcpy(r->w, params, sizeof(float) * r->nparam, hipMemcpyDefault));
  });
}

int mann_router_set_prior(mann_router_t r, mann_ivf_t ivf, double gamma) {
  return guarded([&] {
    MANN_REQUIRE(r != nullptr, ERR_NULL, "null router");
    if (gamma != 0.0) {
      MANN_REQUIRE(ivf != nullptr, ERR_NULL, "null index");
      MANN_REQUIRE(r->c == ivf->nlist && r->d == ivf->d, ERR_DIMENSION,
                   "router does not match the index");
    }
    MANN_HIP(hipSetDevice(r->device));
    router_set_prior(*r, ivf, (float)gamma);
  });
}

int mann_router_grad(mann_router_t r, const float* X, int64_t d, int64_t n, int64_t ldx,
                     const int64_t* labels, int64_t L, float* out_grad, double* out_loss) {
  return guarded([&] {
    MANN_REQUIRE(r && X && labels && out_grad, ERR_NULL, "null pointer");
    MANN_REQUIRE(d == r->d, ERR_DIMENSION, "Expected inputs with " + std::to_string(r->d) + " rows");
    MANN_REQUIRE(n > 0 && L > 0 && ldx >= d, ERR_ARGUMENT, "bad batch");
    ThreadCtx& c = ctx();
    int64_t ld = 0;
    const float* Xd = stage_rows(X, d, n, ldx, "rg_X", &ld);
    const int* lb = stage_labels(labels, n * L, "rg_L");
    float* g = c.buf("rg_grad").as<float>(r->nparam);
    float* loss = c.buf("rg_loss").as<float>(1);
    router_grad(*r, Xd, ld, (int)n, lb, (int)L, g, loss);
    float hl = 0;
    MANN_HIP(hipMemcpyAsync(out_grad, g, sizeof(float) * r->nparam, hipMemcpyDefault, c.stream));
    MANN_HIP(hipMemcpyAsync(&hl, loss, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    MANN_HIP(hipStreamSynchronize(c.stream));
    if (out_loss) *out_loss = hl;
  });
}

int mann_router_train(mann_router_t r, const float* X, int64_t d, int64_t n, int64_t ldx,
                      const int64_t* labels, int64_t L, int64_t epochs, int64_t batch, double lr,
                      uint64_t seed, double* out_loss) {
  return guarded([&] {
    MANN_REQUIRE(r && X && labels, ERR_NULL, "null pointer");
    MANN_REQUIRE(d == r->d, ERR_DIMENSION, "Expected inputs with " + std::to_string(r->d) + " rows");
    MANN_REQUIRE(n > 0 && L > 0 && ldx >= d && epochs > 0 && batch > 0 && batch <= n,
                 ERR_ARGUMENT, "bad training arguments");
    int64_t ld = 0;
    const float* Xd = stage_rows(X, d, n, ldx, "rt_X", &ld);
    const int* lb = stage_labels(labels, n * L, "rt_L");
    double loss = router_train(*r, Xd, ld, n, lb, (int)L, (int)epochs, (int)batch, (float)lr, seed);
    if (out_loss) *out_loss = loss;
    timer_flush();
  });
}

int mann_router_scores(mann_router_t r, const float* Q, int64_t d, int64_t nq, int64_t ldq,
                       float* out_scores) {
  return guarded([&] {
    MANN_REQUIRE(r && Q && out_scores, ERR_NULL, "null pointer");
    MANN_REQUIRE(d == r->d, ERR_DIMENSION, "Expected queries with " + std::to_string(r->d) + " rows");
    if (nq == 0) return;
    int64_t ld = 0;
    const float* Qd = stage_rows(Q, d, nq, ldq, "rs_Q", &ld);
    OutBuf<float> os(out_scores, (size_t)(nq * r->c), "out_scores");
    router_scores(*r, Qd, ld, nq, os.dev);
    os.finish();
    MANN_HIP(hipStreamSynchronize(ctx().stream));
    timer_flush();
  });
}

static void search_routed(mann_ivf_t ivf, mann_router_t r, const float* queries, int64_t d,
                          int64_t nq, int64_t ldq, int64_t k, int64_t nprobe, float mass,
                          int metric, int64_t* out_ids, float* out_dist) {
  MANN_REQUIRE(ivf && r, ERR_NULL, "null handle");
  MANN_REQUIRE(r->c == ivf->nlist && r->d == ivf->d, ERR_DIMENSION,
               "router does not match the index");
  check_metric(metric);
  MANN_REQUIRE((metric == COSINE) == (ivf->normalized != 0), ERR_ARGUMENT,
               "index was built for another metric");
  if (ivf->flat) nprobe = std::max<int64_t>(nprobe, 1);
  MANN_REQUIRE(nprobe > 0, ERR_ARGUMENT, "nprobe must be positive (TopKRouting: k must be positive)");
  QueryOps Q = ivf_queries(ivf, queries, d, nq, ldq);
  if (nq == 0 || k <= 0) {
    last_stats() = SearchStats();
    return;
  }
  MANN_REQUIRE(out_ids && out_dist, ERR_NULL, "null pointer");
  ThreadCtx& c = ctx();
  float* S = c.buf("sr_scores").as<float>((size_t)nq * r->c);
  router_scores(*r, Q.q32, ivf->d16, nq, S, /*negate=*/true);
  OutBuf<int64_t> oi(out_ids, (size_t)(nq * k), "out_ids");
  OutBuf<float> od(out_dist, (size_t)(nq * k), "out_dist");
  ivf_search(*ivf, Q, (int)k, (int)nprobe, metric, S, oi.dev, od.dev, mass);
  oi.finish();
  od.finish();
  timer_flush();
}

int mann_ivf_search_routed(mann_ivf_t ivf, mann_router_t r, const float* queries, int64_t d,
                           int64_t nq, int64_t ldq, int64_t k, int64_t nprobe, int metric,
                           int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    search_routed(ivf, r, queries, d, nq, ldq, k, nprobe, 0.0f, metric, out_ids, out_dist);
  });
}

int mann_ivf_search_routed_mass(mann_ivf_t ivf, mann_router_t r, const float* queries, int64_t d,
                                int64_t nq, int64_t ldq, int64_t k, int64_t nprobe_max,
                                double mass, int metric, int64_t* out_ids, float* out_dist) {
  return guarded([&] {
    MANN_REQUIRE(mass > 0.0 && mass <= 1.0, ERR_ARGUMENT, "mass must be in (0, 1]");
    search_routed(ivf, r, queries, d, nq, ldq, k, nprobe_max, (float)mass, metric, out_ids,
                  out_dist);
  });
}

int mann_ivf_probe_mass(mann_ivf_t ivf, const float* route_dist, int64_t nq, int64_t nprobe_max,
                        double mass, int64_t* out_buckets) {
  return guarded([&] {
    MANN_REQUIRE(mass > 0.0 && mass <= 1.0, ERR_ARGUMENT, "mass must be in (0, 1]");
    probe_out(ivf, route_dist, nq, nprobe_max, (float)mass, out_buckets);
  });
}

int mann_last_probe_count(int64_t* total) {
  return guarded([&] {
    MANN_REQUIRE(total != nullptr, ERR_NULL, "null pointer");
    *total = last_stats().probes;
  });
}

int mann_last_emitted(int64_t* total) {
  return guarded([&] {
    MANN_REQUIRE(total != nullptr, ERR_NULL, "null pointer");
    *total = last_stats().emitted;
  });
}

int mann_tune_set(const char* key, int64_t value) {
  return guarded([&] {
    MANN_REQUIRE(key != nullptr, ERR_NULL, "null key");
    std::string k(key);
    Tuning t = tuning_shared();
    if (k == "scan_cfg" || k == "ivf_cfg" || k == "kmeans_cfg" || k == "ivf_tau_cfg") {
      MANN_REQUIRE(value >= 0 && value < scan_num_cfg(), ERR_ARGUMENT, k + " out of range");
      (k == "scan_cfg"  ? t.scan_cfg
       : k == "ivf_cfg" ? t.ivf_cfg
       : k == "ivf_tau_cfg" ? t.ivf_tau_cfg
                            : t.kmeans_cfg) = (int)value;
    } else if (k == "ivf_first_rows") {
      MANN_REQUIRE(value == 0 || (value >= 16 && value <= (1 << 30)), ERR_ARGUMENT,
                   "ivf_first_rows must be 0 (all) or at least 16");
      t.ivf_first_rows = (int)value;
    } else if (k == "ivf_rows_item") {
      MANN_REQUIRE(value >= 256 && value <= (1 << 20) && value % 256 == 0, ERR_ARGUMENT,
                   "ivf_rows_item must be a multiple of 256 in 256..2^20");
      t.ivf_rows_item = (int)value;
    } else if (k == "xcd_map") {
      MANN_REQUIRE(value >= 0 && value <= 3, ERR_ARGUMENT, "xcd_map must be 0..3");
      t.xcd_map = (int)value;
    } else if (k == "xcd_band") {
      MANN_REQUIRE(value >= 1 && value <= 32, ERR_ARGUMENT, "xcd_band must be in 1..32");
      t.xcd_band = (int)value;
    } else if (k == "scan_cpol") {
      MANN_REQUIRE(value >= 0 && value <= 3, ERR_ARGUMENT, "scan_cpol must be 0..3");
      t.scan_cpol = (int)value;
    } else if (k == "target_groups") {
      MANN_REQUIRE(value >= 1, ERR_ARGUMENT, "target_groups must be positive");
      t.target_groups = (int)value;
    } else if (k == "sample_rank") {
      MANN_REQUIRE(value >= 1 && value <= 64, ERR_ARGUMENT, "sample_rank must be in 1..64");
      t.sample_rank = (int)value;
    } else if (k == "sample_rows") {
      MANN_REQUIRE(value >= 256 && value <= 65536, ERR_ARGUMENT,
                   "sample_rows must be in [256, 65536]");
      t.sample_rows = (int)value;
    } else if (k == "ivf_tau_mode") {
      MANN_REQUIRE(value == 0 || value == 1, ERR_ARGUMENT, "ivf_tau_mode must be 0 or 1");
      t.ivf_tau_mode = (int)value;
    } else if (k == "sample_grp") {
      MANN_REQUIRE(value == 0 || value == 1, ERR_ARGUMENT, "sample_grp must be 0 or 1");
      t.sample_grp = (int)value;
    } else if (k == "ivf_tau_grp") {
      MANN_REQUIRE(value == 0 || value == 1, ERR_ARGUMENT, "ivf_tau_grp must be 0 or 1");
      t.ivf_tau_grp = (int)value;
    } else if (k == "ivf_route_cert") {
      MANN_REQUIRE(value == 0 || value == 1, ERR_ARGUMENT, "ivf_route_cert must be 0 or 1");
      t.ivf_route_cert = (int)value;
    } else if (k == "ivf_route_cfg") {
      MANN_REQUIRE(value >= 0 && value < scan_num_cfg(), ERR_ARGUMENT,
                   "ivf_route_cfg must be a tile configuration index");
      t.ivf_route_cfg = (int)value;
    } else if (k == "tau_rows") {
      MANN_REQUIRE(value >= 16 && value <= 1024, ERR_ARGUMENT, "tau_rows must be in [16, 1024]");
      t.tau_rows = (int)value;
    } else if (k == "shortlist_mult") {
      MANN_REQUIRE(value >= 1 && value <= 16, ERR_ARGUMENT, "shortlist_mult must be in 1..16");
      t.shortlist_mult = (int)value;
    } else if (k == "fold_norms") {
      MANN_REQUIRE(value == 0 || value == 1, ERR_ARGUMENT, "fold_norms must be 0 or 1");
      t.fold_norms = (int)value;
    } else if (k == "emit_per_k") {
      MANN_REQUIRE(value >= 1 && value <= 64, ERR_ARGUMENT, "emit_per_k must be in 1..64");
      t.emit_per_k = (int)value;
    } else if (k == "scan8_cfg") {
      MANN_REQUIRE(value >= 0 && value < scan8_num_cfg(), ERR_ARGUMENT, "scan8_cfg out of range");
      t.scan8_cfg = (int)value;
    } else if (k == "fp8_emit_per_k") {
      MANN_REQUIRE(value >= 1 && value <= 64, ERR_ARGUMENT, "fp8_emit_per_k must be in 1..64");
      t.fp8_emit_per_k = (int)value;
    } else if (k == "sym_join") {
      // the symmetric self-join faulted the GPU on a 580-point self-join and
      // its cause is not identified: it stays off until it is
      MANN_REQUIRE(value == 0, ERR_ARGUMENT,
                   "sym_join is disabled: the symmetric self-join is under investigation");
      t.sym_join = 0;
    } else {
      throw Error{ERR_ARGUMENT, "unknown tuning key " + k};
    }
    tuning_publish(t);
  });
}

int mann_tune_get(const char* key, int64_t* value) {
  return guarded([&] {
    MANN_REQUIRE(key && value, ERR_NULL, "null pointer");
    std::string k(key);
    const Tuning t = tuning_shared();
    if (k == "scan_cfg") *value = t.scan_cfg;
    else if (k == "ivf_cfg") *value = t.ivf_cfg;
    else if (k == "kmeans_cfg") *value = t.kmeans_cfg;
    else if (k == "ivf_tau_cfg") *value = t.ivf_tau_cfg;
    else if (k == "ivf_rows_item") *value = t.ivf_rows_item;
    else if (k == "ivf_first_rows") *value = t.ivf_first_rows;
    else if (k == "xcd_map") *value = t.xcd_map;
    else if (k == "xcd_band") *value = t.xcd_band;
    else if (k == "scan_cpol") *value = t.scan_cpol;
    else if (k == "target_groups") *value = t.target_groups;
    else if (k == "sample_rank") *value = t.sample_rank;
    else if (k == "shortlist_mult") *value = t.shortlist_mult;
    else if (k == "tau_rows") *value = t.tau_rows;
    else if (k == "ivf_tau_mode") *value = t.ivf_tau_mode;
    else if (k == "ivf_tau_grp") *value = t.ivf_tau_grp;
    else if (k == "ivf_route_cert") *value = t.ivf_route_cert;
    else if (k == "ivf_route_cfg") *value = t.ivf_route_cfg;
    else if (k == "sample_grp") *value = t.sample_grp;
    else if (k == "sample_rows") *value = t.sample_rows;
    else if (k == "emit_per_k") *value = t.emit_per_k;
    else if (k == "fold_norms") *value = t.fold_norms;
    else if (k == "scan8_cfg") *value = t.scan8_cfg;
    else if (k == "fp8_emit_per_k") *value = t.fp8_emit_per_k;
    else if (k == "sym_join") *value = t.sym_join;
    else if (k == "last_heap_replays") *value = last_stats().heap_replays;
    else if (k == "last_hnsw_distances") *value = last_stats().hnsw_distances;
    else throw Error{ERR_ARGUMENT, "unknown tuning key " + k};
  });
}

int mann_profile_enable(int on) {
  return guarded([&] { timer().enabled = on != 0; });
}

int mann_profile_reset(void) {
  return guarded([&] {
    timer_flush();
    std::lock_guard<std::mutex> g(timer().mu);
    timer().totals.clear();
  });
}

int mann_profile_query(const char* name, int64_t* count, double* total_ms) {
  return guarded([&] {
    MANN_REQUIRE(name != nullptr, ERR_NULL, "null name");
    timer_flush();
    std::lock_guard<std::mutex> g(timer().mu);
    auto it = timer().totals.find(name);
    if (count) *count = it == timer().totals.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == timer().totals.end() ? 0.0 : it->second.second;
  });
}

int mann_last_stats(int64_t* out4) {
  return guarded([&] {
    MANN_REQUIRE(out4 != nullptr, ERR_NULL, "null pointer");
    const SearchStats& s = last_stats();
    out4[0] = s.reranked;
    out4[1] = s.overflow;
    out4[2] = s.pass2;
    out4[3] = s.exact;
  });
}

}  // extern "C"
