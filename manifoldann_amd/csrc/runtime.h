// Host runtime: per-thread streams and workspaces, device buffers, kernel timing.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "launch.h"

namespace mann {

// A caller-provided block of device memory that scratch buffers are carved
// from in order (no allocation); carving past its end is an ArgumentError.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  size_t used = 0;
  size_t left() const { return size > used ? size - used : 0; }
};

// Grow-only device buffer (or, with an arena, a carve of it).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;
  void* get(size_t need) {
    if (arena) {
      if (p && need <= bytes) return p;
      // the same slack as an allocation below, so both kinds of buffer hold
      // the same bytes
      const size_t want = need + need / 8 + 256;
      const size_t off = (arena->used + 255) / 256 * 256;
      MANN_REQUIRE(off + want <= arena->size, ERR_ARGUMENT,
                   "the allocation-free scratch is exhausted (" + std::to_string(off + want) +
                       " bytes needed, " + std::to_string(arena->size) +
                       " available): split the query batch");
      p = arena->base + off;
      bytes = want;
      arena->used = off + want;
      return p;
    }
    if (need > bytes) {
      if (p) (void)hipFree(p);
      p = nullptr;
      bytes = 0;
      size_t want = need + need / 8 + 256;
      MANN_HIP(hipMalloc(&p, want));
      bytes = want;
    }
    return p;
  }
  template <class T>
  T* as(size_t count) { return (T*)get(count * sizeof(T)); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Set while a search is captured into a HIP graph: the search records what its
// uncaptured tail needs (the exact fallback beyond `fallback_slots` queries,
// the counters) instead of reading the device counters back.
struct CaptureInfo {
  int fallback_slots = 0;
  bool filled = false;
  ExactArgs e{};          // as launched (partials for fallback_slots queries)
  int* flags = nullptr;   // [1]: queries left for the exact fallback
  int* stats = nullptr;   // [kStatSlots][4] selection counters
};

// Per host thread: one stream per device and named scratch buffers.  Every
// exported call runs on the calling thread's stream, so calls from different
// threads never share scratch memory (the re-entrancy contract of
// ann_index.jl:60-70).
struct ThreadCtx {
  int device = -1;
  hipStream_t stream = nullptr;
  std::map<std::string, DevBuf> bufs;
  CaptureInfo* capture = nullptr;
  // set while an allocation-free call runs: buffers are carved from it
  Arena* arena = nullptr;
  std::map<std::string, DevBuf> abufs;
  // set while a captured search is built: its scratch buffers belong to the
  // graph (sized by an eager run, then captured without growing)
  std::map<std::string, DevBuf>* owned = nullptr;
  DevBuf& buf(const std::string& name) {
    if (owned) return (*owned)[name];
    if (arena) {
      DevBuf& b = abufs[name];
      b.arena = arena;
      return b;
    }
    return bufs[name];
  }
};
ThreadCtx& ctx();

// Kernel timing (HIP events on the launching stream), enabled by
// mann_profile_enable().  Records are resolved at the end of each API call.
struct Timer {
  bool enabled = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<long long, double>> totals;  // count, ms
  std::mutex mu;
};
Timer& timer();
// Bracket one launch: begin() returns a token passed to end().
int timer_begin(const char* name, hipStream_t st);
void timer_end(int token, hipStream_t st);
void timer_flush();

bool is_device_ptr(const void* p);

struct ScopedTime {
  int tok;
  hipStream_t st;
  ScopedTime(const char* name, hipStream_t s) : tok(timer_begin(name, s)), st(s) {}
  ~ScopedTime() { timer_end(tok, st); }
};

}  // namespace mann
