// Instantiated hipGraphs of iteration chunks, kept across runs.
//
// Capturing and instantiating a chunk of 16 step launches costs about
// 0.3 ms on MI355X, as much as a short run.  A run that repeats the kernel,
// grid, block size, chunk length and kernel-parameter bytes of an earlier
// run replays the executable captured then.  The key holds the parameter
// bytes themselves, so a cached graph replays exactly the launches a fresh
// capture would record; a parameter block whose bytes differ (another
// schedule, another source functor, reallocated buffers) captures anew.
#ifndef KITES_DETAIL_LAUNCH_CACHE_HPP_
#define KITES_DETAIL_LAUNCH_CACHE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "../common.hpp"
#include "lru.hpp"

namespace kites {
namespace detail {

// A kernel parameter block built in zeroed storage: its padding bytes are
// zero, so blocks with equal members compare equal bytewise.
template <class P>
class param_block {
 public:
  template <class... A>
  explicit param_block(A&&... a) {
    std::memset(raw_, 0, sizeof(raw_));
    p_ = new (raw_) P(std::forward<A>(a)...);
  }
  param_block(param_block const&) = delete;
  param_block& operator=(param_block const&) = delete;
  ~param_block() { p_->~P(); }
  P& get() { return *p_; }
  P const& get() const { return *p_; }

 private:
  alignas(P) unsigned char raw_[sizeof(P)];
  P* p_ = nullptr;
};

// Zeroes a plain parameter struct (padding included) before its members are
// assigned one by one.
template <class T>
inline void zero_params(T& t) {
  static_assert(std::is_trivially_copyable<T>::value, "parameter structs are plain data");
  std::memset(static_cast<void*>(&t), 0, sizeof(T));
}

class launch_cache {
 public:
  static constexpr std::size_t kCapacity = 8;

  launch_cache() : ents_(kCapacity) {}
  launch_cache(launch_cache const&) = delete;
  launch_cache& operator=(launch_cache const&) = delete;
  ~launch_cache() { clear(); }

  // Executable graph of `count` launches of kern<<<grid, block, 0, s>>>(prm),
  // preceded by one launch of pre<<<1, pre_block, 0, s>>>(prm) when `pre` is
  // given, from the cache or captured on `s` now.  Capturing may drop the
  // least recently used executable; callers synchronize their stream after
  // every run, so no dropped executable of an earlier run is still in
  // flight, and the executables returned by the last kCapacity calls stay,
  // so one run can hold one per grid level (three) at once.
  template <class P>
  hipGraphExec_t chunk(void (*kern)(P), unsigned grid, unsigned block, unsigned count,
                       P const& prm, hipStream_t s, void (*pre)(P) = nullptr,
                       unsigned pre_block = 0) {
    static_assert(std::is_trivially_copyable<P>::value, "kernel parameters are copied bytewise");
    std::vector<unsigned char> key(2 * sizeof(kern) + 4 * sizeof(unsigned) + sizeof(P));
    unsigned char* k = key.data();
    std::memcpy(k, &kern, sizeof(kern));
    k += sizeof(kern);
    std::memcpy(k, &pre, sizeof(pre));
    k += sizeof(pre);
    unsigned const dims[4] = {grid, block, count, pre ? pre_block : 0u};
    std::memcpy(k, dims, sizeof(dims));
    k += sizeof(dims);
    std::memcpy(k, static_cast<const void*>(&prm), sizeof(P));
    if (exec_pair* hit = ents_.find(key)) return hit->exec;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    KITES_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (pre) hipLaunchKernelGGL(pre, dim3(1), dim3(pre_block), 0, s, prm);
    for (unsigned i = 0; i < count; ++i)
      hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, prm);
    KITES_HIP_CHECK(hipStreamEndCapture(s, &g));
    hipError_t const ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      KITES_HIP_CHECK(ie);
    }
    ents_.insert(std::move(key), exec_pair{g, x}, release);
    ++builds_;
    return x;
  }

  // Graphs captured so far (cache misses) over the cache's lifetime.
  unsigned builds() const { return builds_; }
  std::size_t size() const { return ents_.size(); }

  void clear() { ents_.clear(release); }

 private:
  struct exec_pair {
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  static void release(exec_pair& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
  }
  byte_lru<exec_pair> ents_;
  unsigned builds_ = 0;
};

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_LAUNCH_CACHE_HPP_
