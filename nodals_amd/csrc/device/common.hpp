// HIP runtime helpers: error checks, owning device buffers, launch shapes.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../host/error.hpp"

namespace nodals {

inline void hipCheck(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess)
    throw Error(kErrHip, std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                             std::to_string(line) + ")");
}
#define NODALS_HIP(x) ::nodals::hipCheck((x), #x, __FILE__, __LINE__)
#define NODALS_LAUNCHED() ::nodals::hipCheck(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Move-only owning device allocation.
template <class T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(std::size_t n) { alloc(n); }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(std::size_t n) {
    release();
    if (n) NODALS_HIP(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const T* h, std::size_t n) {
    if (n != n_) alloc(n);
    if (n) NODALS_HIP(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
  void download(T* h, std::size_t n) const {
    if (n > n_) throw Error(kErrInvalid, "download larger than device buffer");
    if (n) NODALS_HIP(hipMemcpy(h, p_, n * sizeof(T), hipMemcpyDeviceToHost));
  }
  std::vector<T> host() const {
    std::vector<T> v(n_);
    download(v.data(), n_);
    return v;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  std::size_t size() const { return n_; }
  std::size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

// A captured launch sequence replayed with one hipGraphLaunch.  The sequence
// is captured on first use and recaptured when its key (the pointers and
// scalars baked into the kernel arguments) changes.  Captured bodies use
// kernels only (no memset/memcpy nodes).  NODALS_HIP_GRAPHS=0 runs the body
// eagerly instead.
class GraphCache {
 public:
  GraphCache() = default;
  ~GraphCache() { reset(); }
  GraphCache(const GraphCache&) = delete;
  GraphCache& operator=(const GraphCache&) = delete;

  static bool enabled() {
    static const bool on = [] {
      const char* e = std::getenv("NODALS_HIP_GRAPHS");
      return !(e && (std::strcmp(e, "0") == 0 || std::strcmp(e, "off") == 0));
    }();
    return on;
  }

  void reset() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    key_.clear();
  }

  template <class F>
  void run(hipStream_t s, const std::string& key, F&& body) {
    if (!enabled()) {
      body();
      return;
    }
    if (!exec_ || key != key_) {
      reset();
      NODALS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        body();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      NODALS_HIP(hipStreamEndCapture(s, &graph_));
      NODALS_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
      key_ = key;
    }
    NODALS_HIP(hipGraphLaunch(exec_, s));
  }

 private:
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::string key_;
};

// Several captured sequences keyed by their arguments (a V-cycle replayed
// for each (input, output) vector pair of a Krylov basis); at most kMax are
// kept, all dropped when a new key would exceed that.
class GraphSet {
 public:
  template <class F>
  void run(hipStream_t s, const std::string& key, F&& body) {
    auto it = set_.find(key);
    if (it == set_.end()) {
      if (set_.size() >= kMax) set_.clear();
      it = set_.emplace(key, std::make_unique<GraphCache>()).first;
    }
    it->second->run(s, key, std::forward<F>(body));
  }

 private:
  static constexpr std::size_t kMax = 96;
  std::map<std::string, std::unique_ptr<GraphCache>> set_;
};

inline std::string graphKey(std::initializer_list<const void*> ptrs, double scalar = 0.0) {
  std::string k;
  char buf[40];
  for (const void* p : ptrs) {
    std::snprintf(buf, sizeof(buf), "%p;", p);
    k += buf;
  }
  std::snprintf(buf, sizeof(buf), "%.17g", scalar);
  return k + buf;
}

constexpr int kBlock = 256;
inline int blocksFor(long long n, int block = kBlock) { return (int)((n + block - 1) / block); }

// bf16 storage: upper half of an IEEE float, round-to-nearest-even on store.
struct bf16 {
  std::uint16_t bits;
};
__host__ __device__ inline float bf16ToFloat(bf16 h) {
  union {
    std::uint32_t u;
    float f;
  } x;
  x.u = (std::uint32_t)h.bits << 16;
  return x.f;
}
__host__ __device__ inline bf16 floatToBf16(float f) {
  union {
    std::uint32_t u;
    float f;
  } x;
  x.f = f;
  const std::uint32_t lsb = (x.u >> 16) & 1u;
  x.u += 0x7fffu + lsb;
  return bf16{(std::uint16_t)(x.u >> 16)};
}

template <class T>
struct Num {
  __host__ __device__ static double toD(T v) { return (double)v; }
};
template <>
struct Num<bf16> {
  __host__ __device__ static double toD(bf16 v) { return (double)bf16ToFloat(v); }
};

// Load a stored value as the compute type.
template <class C, class T>
__device__ __forceinline__ C ld(const T* p, long long i) {
  return (C)p[i];
}
template <>
__device__ __forceinline__ float ld<float, bf16>(const bf16* p, long long i) {
  return bf16ToFloat(p[i]);
}
template <class T, class C>
__device__ __forceinline__ void st(T* p, long long i, C v) {
  p[i] = (T)v;
}
template <>
__device__ __forceinline__ void st<bf16, float>(bf16* p, long long i, float v) {
  p[i] = floatToBf16(v);
}

}  // namespace nodals
