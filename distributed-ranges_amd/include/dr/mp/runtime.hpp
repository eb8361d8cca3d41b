// dr/mp/runtime.hpp -- the multi-process model's runtime on one process.
//
// The reference runs one MPI rank per device (include/dr/mp/global.hpp:20-197,
// include/dr/detail/communicator.hpp:13-336). Here a "rank" is a host thread
// of a team that shares this process and its GPU: dr::mp::run_ranks(P, fn)
// starts P ranks that each call fn, and dr::mp::default_comm() gives the
// calling rank's communicator. Collectives keep the reference's signatures
// and byte semantics (bcast / gather / gatherv / barrier); the shared address
// space stands in for MPI transport. Outside run_ranks the calling thread is
// rank 0 of a team of one (dr::mp::init()).
//
// Every rank owns a HIP stream on its device (default priority) and a device
// context; kernels of the gemv path are enqueued on it.
#pragma once

#include <hip/hip_runtime_api.h>

#include <barrier>
#include <cstddef>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "dr/detail/drx.hpp"

namespace dr::mp {

namespace __detail {

struct team {
  explicit team(std::size_t n) : size(n), sync(static_cast<std::ptrdiff_t>(n)), slot(n) {}
  std::size_t size;
  std::barrier<> sync;
  std::vector<const void *> slot;  // per-rank published pointer
};

struct rank_context {
  team *t = nullptr;
  std::size_t rank = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialized = false;
  std::unique_ptr<team> owned;  // the team of one outside run_ranks
};

inline rank_context &ctx() {
  thread_local rank_context c;
  return c;
}

}  // namespace __detail

// Communicator of the calling rank (include/dr/detail/communicator.hpp).
class communicator {
public:
  std::size_t rank() const { return __detail::ctx().rank; }
  std::size_t size() const { return team().size; }
  std::size_t prev() const { return (rank() + size() - 1) % size(); }
  std::size_t next() const { return (rank() + 1) % size(); }
  bool first() const { return rank() == 0; }
  bool last() const { return rank() == size() - 1; }

  void barrier() const { team().sync.arrive_and_wait(); }

  // `count` bytes of root's `src` into every rank's `src`.
  void bcast(void *src, std::size_t count, std::size_t root) const {
    auto &t = team();
    t.slot[rank()] = src;
    barrier();
    if (rank() != root && count) std::memcpy(src, t.slot[root], count);
    barrier();
  }

  // `count` bytes from every rank, concatenated in rank order at root's dst.
  void gather(const void *src, void *dst, std::size_t count, std::size_t root) const {
    auto &t = team();
    t.slot[rank()] = src;
    barrier();
    if (rank() == root)
      for (std::size_t r = 0; r < size(); r++)
        if (count)
          std::memcpy(static_cast<char *>(dst) + r * count, t.slot[r], count);
    barrier();
  }

  template <typename T>
  void gather(const T *src, T *dst, std::size_t count, std::size_t root) const {
    gather(static_cast<const void *>(src), static_cast<void *>(dst), count * sizeof(T), root);
  }

  // counts / offsets in bytes, as the reference passes them to MPI_Gatherv_c.
  void gatherv(const void *src, const long long *counts, const long *offsets,
               void *dst, std::size_t root) const {
    auto &t = team();
    t.slot[rank()] = src;
    barrier();
    if (rank() == root)
      for (std::size_t r = 0; r < size(); r++)
        if (counts[r])
          std::memcpy(static_cast<char *>(dst) + offsets[r], t.slot[r],
                      static_cast<std::size_t>(counts[r]));
    barrier();
  }

  // Gather without a copy: every rank publishes `src`; on root the result
  // holds the ranks' pointers in rank order (elsewhere it is empty). The
  // buffers must stay valid and unchanged until every rank has passed a
  // later barrier (the gemv collection fences after the root has read them).
  std::vector<const void *> gather_pointers(const void *src, std::size_t root) const {
    auto &t = team();
    t.slot[rank()] = src;
    barrier();
    std::vector<const void *> out;
    if (rank() == root) out.assign(t.slot.begin(), t.slot.begin() + size());
    barrier();
    return out;
  }

  // Every rank's `count` bytes, concatenated in rank order on every rank.
  void all_gather(const void *src, void *dst, std::size_t count) const {
    auto &t = team();
    t.slot[rank()] = src;
    barrier();
    for (std::size_t r = 0; r < size(); r++)
      if (count) std::memcpy(static_cast<char *>(dst) + r * count, t.slot[r], count);
    barrier();
  }

private:
  static __detail::team &team() {
    auto &c = __detail::ctx();
    if (!c.t) throw std::runtime_error("dr::mp is not initialized (call dr::mp::init)");
    return *c.t;
  }
};

// Rank bring-up (include/dr/mp/global.hpp:130-197): select the device and
// create the rank's stream. Outside run_ranks this makes a team of one.
inline void init(int device = 0) {
  auto &c = __detail::ctx();
  if (c.initialized) return;
  if (!c.t) {
    c.owned = std::make_unique<__detail::team>(1);
    c.t = c.owned.get();
    c.rank = 0;
  }
  c.device = device;
  dr::drx::hip_check(hipSetDevice(device), "hipSetDevice");
  dr::drx::hip_check(hipStreamCreate(&c.stream), "hipStreamCreate");
  c.initialized = true;
}

inline void finalize() {
  auto &c = __detail::ctx();
  if (!c.initialized) return;
  (void)hipStreamSynchronize(c.stream);
  (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.initialized = false;
  if (c.owned) {
    c.t = nullptr;
    c.owned.reset();
  }
}

inline bool finalized() { return !__detail::ctx().initialized; }

inline communicator default_comm() { return communicator{}; }

// The calling rank's stream (the role of dr::mp::sycl_queue()).
inline hipStream_t stream() { return __detail::ctx().stream; }

inline void barrier() { default_comm().barrier(); }

// Run `fn` on P ranks (threads) sharing `device`; rethrows the first
// exception. A rank that throws leaves the others at their next collective,
// so fn should not throw between collectives in tests that expect success.
inline void run_ranks(std::size_t P, const std::function<void()> &fn, int device = 0) {
  __detail::team t(P);
  std::vector<std::thread> th;
  std::exception_ptr err;
  std::mutex mu;
  for (std::size_t r = 0; r < P; r++) {
    th.emplace_back([&, r] {
      auto &c = __detail::ctx();
      c.t = &t;
      c.rank = r;
      try {
        init(device);
        fn();
        finalize();
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
      }
      c.t = nullptr;
    });
  }
  for (auto &x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// The reference's default backend name, kept for source compatibility: here
// a share lives in HBM of the rank's device (hipMalloc), not in an MPI window.
struct MpiBackend {};

// Placeholder for the reference's distribution (halo/granularity) argument;
// the gemv path uses the default (granularity 1, no halo).
struct distribution {
  distribution() = default;
};

namespace __detail {
// Device allocator used by the broadcasted containers (the role of
// dr::mp::__detail::allocator with USM device memory).
template <typename T> struct allocator {
  using value_type = T;
  T *allocate(std::size_t n) {
    T *p = nullptr;
    if (n) dr::drx::hip_check(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T)), "hipMalloc");
    return p;
  }
  void deallocate(T *p, std::size_t) {
    if (p) (void)hipFree(p);
  }
};
}  // namespace __detail

}  // namespace dr::mp
