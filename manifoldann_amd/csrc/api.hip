// C ABI (include/mann.h): argument checks with the reference's exception
// classes, host/device pointer staging, and dispatch to the pipelines.
#include <algorithm>
#include <cstring>
#include <set>
#include <vector>

#include "../../include/mann.h"
#include "core.h"

namespace mann {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

template <class F>
static int guarded(F&& f) {
  try {
    g_err.clear();
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
                                 (int)base->kp, "bfq");
    OutBuf<int64_t> oi(out_ids, (size_t)(nq * k), "out_ids");
    OutBuf<float> od(out_dist, (size_t)(nq * k), "out_dist");
    bf_search(*base, Q, (int)k, metric, 0, oi.dev, od.dev, nullptr);
    oi.finish();
    od.finish();
    timer_flush();
  });
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

int mann_knn_graph_symmetrize(const int64_t* adj, int64_t n, int64_t k, int64_t* out_offsets,
                              int64_t* out_ids, int64_t* out_total, int64_t* out_max_degree) {
  return guarded([&] {
    MANN_REQUIRE(adj && out_offsets && out_ids, ERR_NULL, "null pointer");
    MANN_REQUIRE(n >= 0 && k >= 0, ERR_ARGUMENT, "negative size");
    std::vector<int64_t> deg(n, 0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t r = 0; r < k; ++r) {
        int64_t j = adj[i * k + r];
        MANN_REQUIRE(j >= 1 && j <= n, ERR_ARGUMENT, "neighbour id out of range");
        deg[i]++;
        deg[j - 1]++;
      }
    std::vector<int64_t> off(n + 1, 0), fill(n, 0);
    for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
    std::vector<int64_t> tmp(off[n]);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t r = 0; r < k; ++r) {
        int64_t j = adj[i * k + r] - 1;
        tmp[off[i] + fill[i]++] = j + 1;
        tmp[off[j] + fill[j]++] = i + 1;
      }
    int64_t w = 0, maxd = 0;
    out_offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
      std::sort(tmp.begin() + off[i], tmp.begin() + off[i + 1]);
      int64_t start = w;
      for (int64_t t = off[i]; t < off[i + 1]; ++t)
        if (t == off[i] || tmp[t] != tmp[t - 1]) out_ids[w++] = tmp[t];
      maxd = std::max(maxd, w - start);
      out_offsets[i + 1] = w;
    }
    if (out_total) *out_total = w;
    if (out_max_degree) *out_max_degree = maxd;
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
