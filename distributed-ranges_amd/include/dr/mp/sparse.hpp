// dr/mp/sparse.hpp -- distributed CSR matrix of the multi-process model and
// its two matrix distributions (the gemv plug-in point).
//
//   distributed_sparse_matrix<T, I, Backend, Distr>
//       include/dr/mp/containers/distributed_sparse_matrix.hpp:11-163
//   csr_row_distribution: rank r owns rows [r*seg, (r+1)*seg), seg = ceil(m/P)
//       .../matrix_formats/csr_row_distribution.hpp:13-375
//   csr_eq_distribution: rank r owns entries [r*seg, (r+1)*seg),
//       seg = ceil(nnz/P); its first and last rows may be shared
//       .../matrix_formats/csr_eq_distribution.hpp:13-389
//
// Each rank keeps its share in HBM -- rowptr slice (absolute offsets, no end
// sentinel), colind, vals -- plus a drx plan over it. local_gemv runs
// drx_spmv on the share (beta = 0 into the rank's partial); the partials are
// gathered to the root and assembled with the reference's rules (row: copy
// per rank and right-hand side; eq: zero, add each rank's first row, copy the
// rest).
#pragma once

#include <algorithm>
#include <cassert>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <memory>
#include <ranges>
#include <thread>
#include <vector>

#include "dr/detail/drx.hpp"
#include "dr/detail/index.hpp"
#include "dr/mp/runtime.hpp"
#include "dr/views/csr_matrix_view.hpp"

namespace dr::mp {

namespace __detail {

// Where one rank's share lives (published to the team at construction so
// that any rank can read every share, the role of MPI one-sided gets).
template <typename T, typename I> struct share_desc {
  const T *vals = nullptr;
  const I *colind = nullptr;
  const I *rowptr = nullptr;
  std::int64_t first_row = 0, nrows = 0, base = 0, nnz = 0;
};

// Entries of every share in global order, with global (row, col).
template <typename T, typename I>
std::vector<dr::matrix_entry<T, I>> snapshot(const std::vector<share_desc<T, I>> &shares) {
  std::vector<dr::matrix_entry<T, I>> out;
  for (const auto &s : shares) {
    if (s.nnz == 0) continue;
    std::vector<T> v(s.nnz);
    std::vector<I> c(s.nnz);
    std::vector<I> rp(s.nrows);
    auto st = stream();
    dr::drx::hip_check(hipMemcpyAsync(v.data(), s.vals, s.nnz * sizeof(T), hipMemcpyDeviceToHost, st), "D2H");
    dr::drx::hip_check(hipMemcpyAsync(c.data(), s.colind, s.nnz * sizeof(I), hipMemcpyDeviceToHost, st), "D2H");
    if (s.nrows)
      dr::drx::hip_check(hipMemcpyAsync(rp.data(), s.rowptr, s.nrows * sizeof(I), hipMemcpyDeviceToHost, st), "D2H");
    dr::drx::hip_check(hipStreamSynchronize(st), "sync");
    std::int64_t r = 0;
    for (std::int64_t q = 0; q < s.nnz; q++) {
      while (r + 1 < s.nrows && static_cast<std::int64_t>(rp[r + 1]) - s.base <= q) r++;
      out.emplace_back(dr::index<I>(s.first_row + r, c[q]), v[q]);
    }
  }
  return out;
}

// Column-major partial of `rows` rows (leading dimension ld) on the device.
template <typename T> struct workspace {
  dr::drx::device_buffer<T> buf;
  T *get(std::size_t n) {
    if (buf.size() < n) buf = dr::drx::device_buffer<T>(n);
    return buf.data();
  }
};

template <typename T, typename I>
void upload_share(const T *vals, const I *colind, const I *rowptr, std::size_t nnz,
                  std::size_t rows, dr::drx::device_buffer<T> &v,
                  dr::drx::device_buffer<I> &c, dr::drx::device_buffer<I> &rp) {
  v = dr::drx::device_buffer<T>(std::max<std::size_t>(nnz, 1));
  c = dr::drx::device_buffer<I>(std::max<std::size_t>(nnz, 1));
  rp = dr::drx::device_buffer<I>(std::max<std::size_t>(rows, 1));
  v.upload(vals, nnz, stream());
  c.upload(colind, nnz, stream());
  rp.upload(rowptr, rows, stream());
}

// Copy n partial values into res at out: large copies into contiguous
// memory run on up to 8 threads (a single thread copies a 64 MB partial at
// about 20 GB/s, which was a third of the host part of a gemv call).
template <typename T, typename It> void copy_out(const T *src, std::size_t n, It out) {
  constexpr std::size_t kPar = std::size_t(8) << 20;  // bytes
  if constexpr (std::contiguous_iterator<It>) {
    const std::size_t bytes = n * sizeof(T);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t nt = std::min<std::size_t>({8, hw, bytes / kPar});
    if (nt >= 2) {
      T *dst = std::to_address(out);
      const std::size_t per = (n + nt - 1) / nt;
      std::vector<std::thread> th;
      th.reserve(nt - 1);
      for (std::size_t t = 1; t < nt; t++) {
        const std::size_t b0 = std::min(n, t * per), b1 = std::min(n, b0 + per);
        th.emplace_back([=] { std::memcpy(dst + b0, src + b0, (b1 - b0) * sizeof(T)); });
      }
      std::memcpy(dst, src, std::min(n, per) * sizeof(T));
      for (auto &x : th) x.join();
      return;
    }
  }
  std::copy(src, src + n, out);
}

}  // namespace __detail

// A segment of a distribution: its rank and number of entries.
struct csr_segment {
  std::size_t rank;
  std::size_t size;
};

// ---------------------------------------------------------------------------
// csr_row_distribution
// ---------------------------------------------------------------------------
template <typename T, typename I, class BackendT = MpiBackend>
class csr_row_distribution {
public:
  using value_type = dr::matrix_entry<T, I>;
  using elem_type = T;
  using index_type = I;
  using difference_type = std::ptrdiff_t;

  csr_row_distribution(const csr_row_distribution &) = delete;
  csr_row_distribution &operator=(const csr_row_distribution &) = delete;

  csr_row_distribution(dr::views::csr_matrix_view<T, I> csr_view,
                       distribution dist = distribution(), std::size_t root = 0) {
    (void)dist;
    init(csr_view, root);
  }

  std::size_t get_id_in_segment(std::size_t offset) const {
    assert(offset < nnz_);
    auto it = std::upper_bound(val_offsets_.begin(), val_offsets_.end(), offset) - 1;
    return offset - *it;
  }
  std::size_t get_segment_from_offset(std::size_t offset) const {
    assert(offset < nnz_);
    auto it = std::upper_bound(val_offsets_.begin(), val_offsets_.end(), offset);
    return static_cast<std::size_t>(std::distance(val_offsets_.begin(), it)) - 1;
  }
  const std::vector<csr_segment> &segments() const { return segments_; }
  std::size_t nnz() const { return nnz_; }
  dr::index<std::size_t> shape() const { return shape_; }
  void fence() const { default_comm().barrier(); }
  std::size_t segment_size() const { return segment_size_; }
  // Column panels of this rank's share plan (0: plain CSR). Not in the
  // reference: the share's plan takes the panel layout when its columns
  // scatter over an x larger than one XCD's L2 (drx_plan_auto_panels).
  std::int64_t panels() const { return plan_.panels(); }
  const std::vector<std::size_t> &val_offsets() const { return val_offsets_; }
  const std::vector<std::size_t> &val_sizes() const { return val_sizes_; }

  // y_share = A_share x into the device partial [k][segment_size] (ld = seg).
  void local_gemv(T *partial, const T *x, std::size_t k) const {
    if (rows_ == 0) return;
    dr::drx::check(drx_spmv(plan_.get(), dr::drx::dtype_code<T>(), cols_.data(), vals_.data(),
                            dr::drx::dtype_code<T>(), x, static_cast<std::int64_t>(shape_[1]),
                            dr::drx::dtype_code<T>(), partial,
                            static_cast<std::int64_t>(segment_size_),
                            static_cast<std::int64_t>(k), 1.0, 0.0, stream()),
                   "drx_spmv");
  }

  // csr_row_distribution.hpp:136-153. The partial lives in pinned host
  // memory kept across calls (the reference allocates and zero-fills it per
  // call); ranks without rows contribute nothing the root reads.
  template <typename C>
  void local_gemv_and_collect(std::size_t root, C &res, T *x, std::size_t k) const {
    auto comm = default_comm();
    if (comm.rank() == root) assert(std::ranges::size(res) == shape_[0] * k);
    const std::size_t n_part = segment_size_ * k;
    T *partial = hpart_.get(std::max<std::size_t>(n_part, 1));
    if (rows_ > 0) {
      T *d = ws_.get(std::max<std::size_t>(n_part, 1));
      local_gemv(d, x, k);
      dr::drx::hip_check(hipMemcpyAsync(partial, d, n_part * sizeof(T),
                                        hipMemcpyDeviceToHost, stream()), "D2H");
      dr::drx::hip_check(hipStreamSynchronize(stream()), "sync");
    }
    gather_gemv_vector(root, res, partial, k);
    fence();
  }

  std::vector<value_type> entries() const { return __detail::snapshot(shares_); }

private:
  // csr_row_distribution.hpp:158-204: Gather of equal-size partials, then the
  // root copies min(seg, m - j*seg) rows of every rank and right-hand side.
  // The root reads each rank's partial in place (the shared address space
  // is the transport), so there is no intermediate gather buffer.
  template <typename C>
  void gather_gemv_vector(std::size_t root, C &res, const T *partial, std::size_t k) const {
    auto comm = default_comm();
    const std::size_t P = comm.size(), m = shape_[0];
    const auto parts = comm.gather_pointers(partial, root);
    if (comm.rank() != root) return;
    auto out = std::ranges::begin(res);
    for (std::size_t j = 0; j < P; j++) {
      if (j * segment_size_ >= m) break;
      const std::size_t css = std::min(segment_size_, m - j * segment_size_);
      for (std::size_t i = 0; i < k; i++) {
        const T *piece = static_cast<const T *>(parts[j]) + i * segment_size_;
        __detail::copy_out(piece, css, out + m * i + j * segment_size_);
      }
    }
  }

  // csr_row_distribution.hpp:205-318: sizes from the root, row segments of
  // ceil(m/P), each rank's entry range from the rowptr at its first row.
  void init(dr::views::csr_matrix_view<T, I> csr_view, std::size_t root) {
    auto comm = default_comm();
    const std::size_t rank = comm.rank(), P = comm.size();
    std::size_t info[3] = {0, 0, 0};
    const void *arrays[3] = {nullptr, nullptr, nullptr};
    if (rank == root) {
      info[0] = csr_view.size();
      info[1] = csr_view.shape()[0];
      info[2] = csr_view.shape()[1];
      arrays[0] = csr_view.values_data();
      arrays[1] = csr_view.rowptr_data();
      arrays[2] = csr_view.colind_data();
    }
    comm.bcast(info, sizeof(info), root);
    comm.bcast(arrays, sizeof(arrays), root);
    nnz_ = info[0];
    shape_ = {info[1], info[2]};
    const auto *vals = static_cast<const T *>(arrays[0]);
    const auto *rowptr = static_cast<const I *>(arrays[1]);
    const auto *colind = static_cast<const I *>(arrays[2]);
    const std::size_t m = shape_[0];
    segment_size_ = (m + P - 1) / P;
    std::vector<std::size_t> info2(2 * P);
    if (rank == root) {
      for (std::size_t i = 0; i < P; i++) {
        const std::size_t first = i * segment_size_;
        if (first > m) {
          info2[i] = nnz_;
          info2[P + i] = 0;
          continue;
        }
        const std::size_t lower = static_cast<std::size_t>(rowptr[first]);
        std::size_t higher = nnz_;
        if ((i + 1) * segment_size_ < m)
          higher = static_cast<std::size_t>(rowptr[(i + 1) * segment_size_]);
        info2[i] = lower;
        info2[P + i] = higher - lower;
      }
    }
    comm.bcast(info2.data(), info2.size() * sizeof(std::size_t), root);
    val_offsets_.assign(info2.begin(), info2.begin() + P);
    val_sizes_.assign(info2.begin() + P, info2.end());
    for (std::size_t i = 0; i < P; i++) segments_.push_back({i, val_sizes_[i]});

    rows_ = m > segment_size_ * rank ? std::min(segment_size_, m - segment_size_ * rank) : 0;
    const std::size_t my_nnz = val_sizes_[rank], off = val_offsets_[rank];
    __detail::upload_share(vals + std::min(off, nnz_), colind + std::min(off, nnz_),
                           rowptr + std::min(segment_size_ * rank, m), my_nnz, rows_,
                           vals_, cols_, rowptr_);
    if (rows_ > 0) {
      plan_ = dr::drx::plan(rowptr_.data(), static_cast<std::int64_t>(off),
                            static_cast<std::int64_t>(rows_),
                            static_cast<std::int64_t>(shape_[1]),
                            static_cast<std::int64_t>(my_nnz), false, stream());
      plan_.auto_panels(dr::drx::dtype_code<T>(), cols_.data(), vals_.data(), stream());
    }
    __detail::share_desc<T, I> me{vals_.data(), cols_.data(), rowptr_.data(),
                                  static_cast<std::int64_t>(segment_size_ * rank),
                                  static_cast<std::int64_t>(rows_),
                                  static_cast<std::int64_t>(off),
                                  static_cast<std::int64_t>(my_nnz)};
    shares_.resize(P);
    comm.all_gather(&me, shares_.data(), sizeof(me));
    fence();
  }

  std::size_t segment_size_ = 0, rows_ = 0, nnz_ = 0;
  dr::index<std::size_t> shape_;
  std::vector<std::size_t> val_offsets_, val_sizes_;
  std::vector<csr_segment> segments_;
  dr::drx::device_buffer<T> vals_;
  dr::drx::device_buffer<I> cols_, rowptr_;
  dr::drx::plan plan_;
  std::vector<__detail::share_desc<T, I>> shares_;
  mutable __detail::workspace<T> ws_;
  mutable dr::drx::pinned_buffer<T> hpart_;  // the partial on the host
};

// ---------------------------------------------------------------------------
// csr_eq_distribution
// ---------------------------------------------------------------------------
template <typename T, typename I, class BackendT = MpiBackend>
class csr_eq_distribution {
public:
  using value_type = dr::matrix_entry<T, I>;
  using elem_type = T;
  using index_type = I;
  using difference_type = std::ptrdiff_t;

  csr_eq_distribution(const csr_eq_distribution &) = delete;
  csr_eq_distribution &operator=(const csr_eq_distribution &) = delete;

  csr_eq_distribution(dr::views::csr_matrix_view<T, I> csr_view,
                      distribution dist = distribution(), std::size_t root = 0) {
    (void)dist;
    init(csr_view, root);
  }

  std::size_t get_id_in_segment(std::size_t offset) const { return offset % segment_size_; }
  std::size_t get_segment_from_offset(std::size_t offset) const { return offset / segment_size_; }
  const std::vector<csr_segment> &segments() const { return segments_; }
  std::size_t nnz() const { return nnz_; }
  dr::index<std::size_t> shape() const { return shape_; }
  void fence() const { default_comm().barrier(); }
  std::size_t segment_size() const { return segment_size_; }
  // Column panels of this rank's share plan (0: plain CSR). Not in the
  // reference: the share's plan takes the panel layout when its columns
  // scatter over an x larger than one XCD's L2 (drx_plan_auto_panels).
  std::int64_t panels() const { return plan_.panels(); }
  const std::vector<std::size_t> &row_offsets() const { return row_offsets_; }
  const std::vector<std::size_t> &row_sizes() const { return row_sizes_; }

  // Partial [k][row_sizes[rank]] of the rank's entries; the first and last
  // rows of the share hold partial sums of rows shared with the neighbours.
  void local_gemv(T *partial, const T *x, std::size_t k) const {
    auto comm = default_comm();
    const std::size_t rank = comm.rank(), rows = row_sizes_[rank];
    if (nnz_ <= segment_size_ * rank) {
      if (rows * k)
        dr::drx::hip_check(hipMemsetAsync(partial, 0, rows * k * sizeof(T), stream()), "memset");
      return;
    }
    dr::drx::check(drx_spmv(plan_.get(), dr::drx::dtype_code<T>(), cols_.data(), vals_.data(),
                            dr::drx::dtype_code<T>(), x, static_cast<std::int64_t>(shape_[1]),
                            dr::drx::dtype_code<T>(), partial, static_cast<std::int64_t>(rows),
                            static_cast<std::int64_t>(k), 1.0, 0.0, stream()),
                   "drx_spmv");
  }

  // csr_eq_distribution.hpp:133-153, with the partial in pinned host memory
  // kept across calls.
  template <typename C>
  void local_gemv_and_collect(std::size_t root, C &res, T *x, std::size_t k) const {
    auto comm = default_comm();
    if (comm.rank() == root) assert(std::ranges::size(res) == shape_[0] * k);
    const std::size_t n_part = row_sizes_[comm.rank()] * k;
    T *partial = hpart_.get(std::max<std::size_t>(n_part, 1));
    if (n_part) {
      T *d = ws_.get(n_part);
      local_gemv(d, x, k);
      dr::drx::hip_check(hipMemcpyAsync(partial, d, n_part * sizeof(T),
                                        hipMemcpyDeviceToHost, stream()), "D2H");
      dr::drx::hip_check(hipStreamSynchronize(stream()), "sync");
    }
    gather_gemv_vector(root, res, partial, k);
    fence();
  }

  std::vector<value_type> entries() const { return __detail::snapshot(shares_); }

private:
  // csr_eq_distribution.hpp:158-216: Gatherv of [k][row_sizes[i]] partials;
  // the root zeroes res, adds every rank's first row (shared with the
  // previous rank) and copies its other rows. Here the root reads each
  // rank's partial in place, and zeroes only what no copy writes first: the
  // rows before a rank's first row that no earlier rank covered, that first
  // row itself when the previous rank did not write it, and the rows after
  // the last rank's. Every element ends with the bits the reference's
  // zero-add-copy order gives (0 + a == a; boundary sums in rank order).
  template <typename C>
  void gather_gemv_vector(std::size_t root, C &res, const T *partial, std::size_t k) const {
    auto comm = default_comm();
    const std::size_t P = comm.size(), m = shape_[0];
    const auto parts = comm.gather_pointers(partial, root);
    if (comm.rank() != root) return;
    auto out = std::ranges::begin(res);
    for (std::size_t kk = 0; kk < k; kk++) {
      auto col = out + kk * m;
      std::size_t written = 0;  // rows [0, written) of this column hold their value so far
      for (std::size_t i = 0; i < P; i++) {
        const std::size_t first = std::min(row_offsets_[i], m);
        const std::size_t last = std::min(first + row_sizes_[i], m);
        if (first >= last) continue;
        const T *piece = static_cast<const T *>(parts[i]) + kk * row_sizes_[i];
        if (written <= first) std::fill(col + written, col + first + 1, T(0));
        col[first] += piece[0];
        __detail::copy_out(piece + 1, last - first - 1, col + first + 1);
        written = std::max(written, last);
      }
      if (written < m) std::fill(col + written, col + m, T(0));
    }
  }

  // csr_eq_distribution.hpp:220-333: nnz segments of ceil(nnz/P); each
  // rank's rows from upper_bound on the root's rowptr.
  void init(dr::views::csr_matrix_view<T, I> csr_view, std::size_t root) {
    auto comm = default_comm();
    const std::size_t rank = comm.rank(), P = comm.size();
    std::size_t info[3] = {0, 0, 0};
    const void *arrays[3] = {nullptr, nullptr, nullptr};
    if (rank == root) {
      info[0] = csr_view.size();
      info[1] = csr_view.shape()[0];
      info[2] = csr_view.shape()[1];
      arrays[0] = csr_view.values_data();
      arrays[1] = csr_view.rowptr_data();
      arrays[2] = csr_view.colind_data();
    }
    comm.bcast(info, sizeof(info), root);
    comm.bcast(arrays, sizeof(arrays), root);
    nnz_ = info[0];
    shape_ = {info[1], info[2]};
    const auto *vals = static_cast<const T *>(arrays[0]);
    const auto *rowptr = static_cast<const I *>(arrays[1]);
    const auto *colind = static_cast<const I *>(arrays[2]);
    const std::size_t m = shape_[0];
    segment_size_ = (nnz_ + P - 1) / P;
    std::vector<std::size_t> info2(2 * P + 1, 0);
    if (rank == root) {
      auto ub = [&](std::size_t v) {
        return static_cast<std::size_t>(
            std::upper_bound(rowptr, rowptr + m, v,
                             [](std::size_t a, const I &e) { return a < static_cast<std::size_t>(e); }) -
            rowptr);
      };
      std::size_t total = 0;
      for (std::size_t i = 0; i < P; i++) {
        const std::size_t first_index = i * segment_size_;
        const std::size_t last_index = (i + 1) * segment_size_ - 1;
        const std::size_t lower = ub(first_index) - 1;
        const std::size_t higher = ub(last_index);
        info2[i] = lower;
        info2[P + i] = higher - lower;
        total += higher - lower;
      }
      info2[2 * P] = total;
    }
    comm.bcast(info2.data(), info2.size() * sizeof(std::size_t), root);
    row_offsets_.assign(info2.begin(), info2.begin() + P);
    row_sizes_.assign(info2.begin() + P, info2.begin() + 2 * P);
    max_row_size_ = info2[2 * P];
    for (std::size_t i = 0; i < nnz_; i += segment_size_)
      segments_.push_back({segments_.size(), std::min(segment_size_, nnz_ - i)});

    const std::size_t lo = std::min(segment_size_ * rank, nnz_);
    const std::size_t my_nnz = std::min(segment_size_ * (rank + 1), nnz_) - lo;
    const std::size_t rows = row_sizes_[rank];
    const std::size_t rfirst = std::min(row_offsets_[rank], m);
    const std::size_t rows_avail = std::min(rows, m - rfirst);
    __detail::upload_share(vals + lo, colind + lo, rowptr + rfirst, my_nnz, rows_avail,
                           vals_, cols_, rowptr_);
    if (my_nnz > 0) {
      plan_ = dr::drx::plan(rowptr_.data(), static_cast<std::int64_t>(lo),
                            static_cast<std::int64_t>(rows),
                            static_cast<std::int64_t>(shape_[1]),
                            static_cast<std::int64_t>(my_nnz), false, stream());
      plan_.auto_panels(dr::drx::dtype_code<T>(), cols_.data(), vals_.data(), stream());
    }
    __detail::share_desc<T, I> me{vals_.data(), cols_.data(), rowptr_.data(),
                                  static_cast<std::int64_t>(rfirst),
                                  static_cast<std::int64_t>(rows_avail),
                                  static_cast<std::int64_t>(lo),
                                  static_cast<std::int64_t>(my_nnz)};
    shares_.resize(P);
    comm.all_gather(&me, shares_.data(), sizeof(me));
    fence();
  }

  std::size_t segment_size_ = 0, nnz_ = 0, max_row_size_ = 0;
  dr::index<std::size_t> shape_;
  std::vector<std::size_t> row_offsets_, row_sizes_;
  std::vector<csr_segment> segments_;
  dr::drx::device_buffer<T> vals_;
  dr::drx::device_buffer<I> cols_, rowptr_;
  dr::drx::plan plan_;
  std::vector<__detail::share_desc<T, I>> shares_;
  mutable __detail::workspace<T> ws_;
  mutable dr::drx::pinned_buffer<T> hpart_;  // the partial on the host
};

// ---------------------------------------------------------------------------
// distributed_sparse_matrix and its plug-in concepts
// ---------------------------------------------------------------------------
template <typename T>
concept matrix_distibution = requires(T t) {
  { t.fence() } -> std::same_as<void>;
  { t.shape().first } -> std::convertible_to<std::size_t>;
  { t.shape().second } -> std::convertible_to<std::size_t>;
  { t.nnz() } -> std::same_as<std::size_t>;
  { t.get_segment_from_offset(int()) } -> std::same_as<std::size_t>;
  { t.get_id_in_segment(int()) } -> std::same_as<std::size_t>;
};

template <typename T>
concept vector_multiplicable =
    requires(T t, std::vector<typename T::elem_type> res, typename T::elem_type *input) {
      t.local_gemv_and_collect(int(), res, input, 1);
    };

template <typename T, typename I, class BackendT = MpiBackend,
          class MatrixDistrT = csr_eq_distribution<T, I, BackendT>>
  requires(matrix_distibution<MatrixDistrT>)
class distributed_sparse_matrix {
public:
  using value_type = dr::matrix_entry<T, I>;
  using elem_type = T;
  using index_type = I;
  using key_type = dr::index<I>;
  using size_type = std::size_t;
  using difference_type = std::ptrdiff_t;
  using backend_type = BackendT;
  using iterator = typename std::vector<value_type>::const_iterator;

  distributed_sparse_matrix(const distributed_sparse_matrix &) = delete;
  distributed_sparse_matrix &operator=(const distributed_sparse_matrix &) = delete;

  distributed_sparse_matrix(dr::views::csr_matrix_view<T, I> csr_view, std::size_t root = 0,
                            dr::mp::distribution dist = dr::mp::distribution())
      : distribution_(csr_view, dist, root) {}

  // Entries in global order (read from every rank's share in HBM).
  iterator begin() const { return entries().begin(); }
  iterator end() const { return entries().end(); }
  std::size_t size() const { return distribution_.nnz(); }
  auto shape() const { return distribution_.shape(); }
  value_type operator[](difference_type n) const { return entries()[n]; }
  const auto &segments() const { return distribution_.segments(); }
  void fence() { distribution_.fence(); }
  const MatrixDistrT &matrix_distribution() const { return distribution_; }

  template <typename C>
    requires(vector_multiplicable<MatrixDistrT>)
  void local_gemv_and_collect(std::size_t root, C &res, T *vals, std::size_t val_width) const {
    distribution_.local_gemv_and_collect(root, res, vals, val_width);
  }

private:
  const std::vector<value_type> &entries() const {
    if (!snapshot_)
      snapshot_ = std::make_shared<std::vector<value_type>>(distribution_.entries());
    return *snapshot_;
  }
  MatrixDistrT distribution_;
  mutable std::shared_ptr<std::vector<value_type>> snapshot_;
};

}  // namespace dr::mp
