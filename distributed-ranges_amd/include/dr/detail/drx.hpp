// dr/detail/drx.hpp -- glue between the host interface and the C ABI
// (include/drx_spmv.h): dtype codes, status checks, owning device buffers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "drx_spmv.h"

namespace dr::drx {

template <typename T> constexpr int dtype_code() {
  if constexpr (std::is_same_v<T, double>)
    return DRX_F64;
  else if constexpr (std::is_same_v<T, float>)
    return DRX_F32;
  else
    static_assert(sizeof(T) == 0, "the gemv host interface takes float or double");
}

template <typename I> constexpr int itype_code() {
  static_assert(std::is_integral_v<I> && (sizeof(I) == 4 || sizeof(I) == 8),
                "index types are 32- or 64-bit integers");
  return sizeof(I) == 4 ? DRX_I32 : DRX_I64;
}

inline void check(int status, const char *what) {
  if (status != DRX_OK)
    throw std::runtime_error(std::string(what) + ": " + drx_last_error());
}

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device allocation of n elements (hipMalloc / hipFree).
template <typename T> class device_buffer {
public:
  device_buffer() = default;
  explicit device_buffer(std::size_t n) : n_(n) {
    if (n_) hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n_ * sizeof(T)), "hipMalloc");
  }
  device_buffer(const device_buffer &) = delete;
  device_buffer &operator=(const device_buffer &) = delete;
  device_buffer(device_buffer &&o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  device_buffer &operator=(device_buffer &&o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~device_buffer() { reset(); }
  void reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *data() const noexcept { return p_; }
  std::size_t size() const noexcept { return n_; }
  void upload(const T *host, std::size_t n, hipStream_t s = nullptr) {
    if (n) {
      hip_check(hipMemcpyAsync(p_, host, n * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
      hip_check(hipStreamSynchronize(s), "sync");
    }
  }
  void download(T *host, std::size_t n, hipStream_t s = nullptr) const {
    if (n) {
      hip_check(hipMemcpyAsync(host, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "D2H");
      hip_check(hipStreamSynchronize(s), "sync");
    }
  }

private:
  T *p_ = nullptr;
  std::size_t n_ = 0;
};

// Pinned host memory (hipHostMalloc), grown on demand and kept: the staging
// buffer of a rank's partial result, copied from the device at full link
// speed and read in place by the root's assembly.
template <typename T> class pinned_buffer {
public:
  pinned_buffer() = default;
  pinned_buffer(const pinned_buffer &) = delete;
  pinned_buffer &operator=(const pinned_buffer &) = delete;
  ~pinned_buffer() { reset(); }
  T *get(std::size_t n) {
    if (n > n_) {
      reset();
      hip_check(hipHostMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T), hipHostMallocDefault),
                "hipHostMalloc");
      n_ = n;
    }
    return p_;
  }
  void reset() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }

private:
  T *p_ = nullptr;
  std::size_t n_ = 0;
};

// A plan over one CSR share (drx_plan_create / drx_plan_destroy).
class plan {
public:
  plan() = default;
  template <typename I>
  plan(const I *rowptr, std::int64_t base, std::int64_t nrows,
       std::int64_t ncols, std::int64_t nnz, bool sentinel, hipStream_t s) {
    check(drx_plan_create(&h_, itype_code<I>(), rowptr, base, nrows, ncols, nnz,
                          sentinel ? 1 : 0, 0, s),
          "drx_plan_create");
  }
  plan(const plan &) = delete;
  plan &operator=(const plan &) = delete;
  plan(plan &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  plan &operator=(plan &&o) noexcept {
    if (this != &o) {
      if (h_) drx_plan_destroy(h_);
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }
  ~plan() {
    if (h_) drx_plan_destroy(h_);
  }
  drx_plan get() const noexcept { return h_; }
  // Column-panel layout when it pays (drx_plan_auto_panels); true if built.
  bool auto_panels(int vtype, const void *colind, const void *vals, hipStream_t s) {
    int built = 0;
    if (h_) check(drx_plan_auto_panels(h_, vtype, colind, vals, s, &built), "drx_plan_auto_panels");
    return built != 0;
  }
  // Column panels of the plan (0 without the layout).
  std::int64_t panels() const {
    std::int64_t n = 0;
    if (h_) check(drx_plan_panels(h_, &n, nullptr, nullptr), "drx_plan_panels");
    return n;
  }

private:
  drx_plan h_ = nullptr;
};

}  // namespace dr::drx
