// dr/sp.hpp -- the single-process model's sparse gemv interface on MI355X.
//
//   dr::sp::init / nprocs / devices            include/dr/sp/init.hpp:48-84
//   block_cyclic / row_cyclic / tile::div      .../containers/matrix_partition.hpp:16-94
//   distributed_vector, duplicated_vector      .../distributed_vector.hpp,
//                                              .../containers/duplicated_vector.hpp:12-46
//   sparse_matrix (tiles in HBM, tile())       .../containers/sparse_matrix.hpp:135-425
//   create_distributed, mmread                  .../util/matrix_io.hpp:20-84
//   gemv(c, a, b[, scratch])                    .../algorithms/matrix/gemv.hpp:78-124
//   flat_gemv, gemv_square, gemv_square_copy    .../algorithms/matrix/gemv.hpp:19-76,126-207
//   __detail::local_gemv (per tile, c += A b)   .../algorithms/matrix/local_gemv.hpp:83-98
//   detail::factor, partition_matmul            .../containers/detail.hpp:15-24,
//                                              .../containers/matrix_partition.hpp:96-117
//   dense_matrix_view, distributed_dense_matrix .../views/dense_matrix_view.hpp,
//                                              .../containers/distributed_dense_matrix.hpp:136-331
//   gemm, gemm_inplace, gemm_buffered(_async)   .../algorithms/matrix/gemm.hpp:12-245
//   __detail::local_gemm (per tile, c += a b)   .../algorithms/matrix/local_gemm.hpp:15-60
//
// A "device" of this model is a logical device: an entry of the list given
// to init, each with its own HIP stream. Listing the same GPU several times
// splits work into several tiles on it (the reference's tests do the same
// with --devicesCount). Tile kernels run through drx_spmv_csr with
// alpha = beta = 1: like oneMKL's sparse::gemv as the reference calls it,
// they accumulate into c.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "dr/detail/drx.hpp"
#include "dr/detail/generate_csr.hpp"
#include "dr/detail/index.hpp"
#include "dr/detail/matrix_io.hpp"
#include "dr/views/csr_matrix_view.hpp"

namespace dr::sp {

// ---------------------------------------------------------------------------
// Runtime
// ---------------------------------------------------------------------------
namespace __detail {
struct runtime {
  std::vector<int> devices;
  std::vector<hipStream_t> streams;
};
inline runtime &rt() {
  static runtime r;
  return r;
}
}  // namespace __detail

inline void finalize() {
  auto &r = __detail::rt();
  for (auto s : r.streams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  r.streams.clear();
  r.devices.clear();
}

// One logical device per entry (repeat an id to place several on one GPU).
inline void init(const std::vector<int> &devices) {
  finalize();
  auto &r = __detail::rt();
  r.devices = devices.empty() ? std::vector<int>{0} : devices;
  for (int d : r.devices) {
    dr::drx::hip_check(hipSetDevice(d), "hipSetDevice");
    hipStream_t s;
    dr::drx::hip_check(hipStreamCreate(&s), "hipStreamCreate");
    r.streams.push_back(s);
  }
  dr::drx::hip_check(hipSetDevice(r.devices[0]), "hipSetDevice");
}

// `count` logical devices on GPU 0.
inline void init(std::size_t count = 1) { init(std::vector<int>(count ? count : 1, 0)); }

inline std::size_t nprocs() { return __detail::rt().devices.size(); }
inline const std::vector<int> &devices() { return __detail::rt().devices; }

namespace __detail {
inline hipStream_t queue(std::size_t rank) { return rt().streams.at(rank); }
inline void wait_all() {
  for (auto s : rt().streams) dr::drx::hip_check(hipStreamSynchronize(s), "sync");
}
}  // namespace __detail

// ---------------------------------------------------------------------------
// Partitions
// ---------------------------------------------------------------------------
namespace tile {
inline constexpr std::size_t div = std::numeric_limits<std::size_t>::max();
}

class matrix_partition {
public:
  virtual std::size_t tile_rank(dr::index<> matrix_shape, dr::index<> tile_id) const = 0;
  virtual dr::index<> grid_shape(dr::index<> matrix_shape) const = 0;
  virtual dr::index<> tile_shape(dr::index<> matrix_shape) const = 0;
  virtual std::unique_ptr<matrix_partition> clone() const = 0;
  virtual ~matrix_partition() = default;
};

namespace detail {
// Factor pair (p, q) of n with q the largest divisor of n at most sqrt(n).
inline dr::index<> factor(std::size_t n) {
  std::size_t q = static_cast<std::size_t>(std::sqrt(static_cast<double>(n)));
  while (q > 1 && n % q != 0) q--;
  q = std::max<std::size_t>(q, 1);
  return dr::index<>(n / q, q);
}
}  // namespace detail

class block_cyclic final : public matrix_partition {
public:
  block_cyclic(dr::index<> tile_shape = {tile::div, tile::div},
               dr::index<> grid_shape = detail::factor(nprocs()))
      : tile_shape_(tile_shape), grid_shape_(grid_shape) {}

  std::size_t tile_rank(dr::index<>, dr::index<> tile_id) const override {
    const std::size_t gi = tile_id[0] % grid_shape_[0], gj = tile_id[1] % grid_shape_[1];
    return (gi * grid_shape_[1] + gj) % std::max<std::size_t>(nprocs(), 1);
  }
  dr::index<> grid_shape(dr::index<> s) const override {
    auto ts = tile_shape(s);
    return dr::index<>((s[0] + ts[0] - 1) / ts[0], (s[1] + ts[1] - 1) / ts[1]);
  }
  dr::index<> tile_shape(dr::index<> s) const override {
    dr::index<> t = tile_shape_;
    for (std::size_t d = 0; d < 2; d++)
      if (t[d] == tile::div) t[d] = std::max<std::size_t>((s[d] + grid_shape_[d] - 1) / grid_shape_[d], 1);
    return t;
  }
  std::unique_ptr<matrix_partition> clone() const override {
    return std::make_unique<block_cyclic>(*this);
  }

private:
  dr::index<> tile_shape_;
  dr::index<> grid_shape_;
};

inline block_cyclic row_cyclic() { return block_cyclic({tile::div, tile::div}, {nprocs(), 1}); }

// ---------------------------------------------------------------------------
// Vectors
// ---------------------------------------------------------------------------
template <typename T> class device_vector {
public:
  device_vector() = default;
  device_vector(std::size_t n, std::size_t rank) : buf_(n), rank_(rank) {}
  T *data() const { return buf_.data(); }
  std::size_t size() const { return buf_.size(); }
  std::size_t rank() const { return rank_; }

private:
  dr::drx::device_buffer<T> buf_;
  std::size_t rank_ = 0;
};

template <typename T> struct vector_segment {
  T *data;
  std::size_t size;
  std::size_t rank;
};

template <typename T> class distributed_vector {
public:
  using value_type = T;
  class iterator {
  public:
    using iterator_category = std::random_access_iterator_tag;
    using value_type = T;
    using difference_type = std::ptrdiff_t;
    using pointer = T *;
    using reference = T &;
    iterator() = default;
    iterator(const distributed_vector *v, std::size_t i) : v_(v), i_(i) {}
    iterator &operator++() { ++i_; return *this; }
    iterator operator+(difference_type n) const { return iterator(v_, i_ + n); }
    difference_type operator-(const iterator &o) const { return difference_type(i_) - difference_type(o.i_); }
    bool operator==(const iterator &o) const { return i_ == o.i_; }
    const distributed_vector *vec() const { return v_; }
    std::size_t pos() const { return i_; }

  private:
    const distributed_vector *v_ = nullptr;
    std::size_t i_ = 0;
  };

  explicit distributed_vector(std::size_t count) : size_(count) {
    const std::size_t P = nprocs();
    seg_ = std::max<std::size_t>((count + P - 1) / P, 1);
    for (std::size_t r = 0; r * seg_ < count; r++)
      store_.emplace_back(std::min(seg_, count - r * seg_), r);
  }
  distributed_vector(std::size_t count, T value) : distributed_vector(count) {
    std::vector<T> fill(seg_, value);
    for (auto &s : store_) dr::drx::hip_check(hipMemcpy(s.data(), fill.data(), s.size() * sizeof(T), hipMemcpyHostToDevice), "fill");
  }
  std::size_t size() const { return size_; }
  std::size_t segment_size() const { return seg_; }
  std::vector<vector_segment<T>> segments() const {
    std::vector<vector_segment<T>> out;
    for (auto &s : store_) out.push_back({s.data(), s.size(), s.rank()});
    return out;
  }
  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, size_); }

private:
  std::size_t size_ = 0, seg_ = 1;
  std::vector<device_vector<T>> store_;
};

// Device -> host copy of [first, last) of a distributed vector.
template <typename T, typename OutIt>
OutIt copy(typename distributed_vector<T>::iterator first,
           typename distributed_vector<T>::iterator last, OutIt out) {
  auto segs = first.vec()->segments();
  const std::size_t seg = first.vec()->segment_size();
  std::vector<T> tmp;
  for (std::size_t i = first.pos(); i < last.pos();) {
    const std::size_t s = i / seg, off = i % seg;
    const std::size_t n = std::min(segs[s].size - off, last.pos() - i);
    tmp.resize(n);
    dr::drx::hip_check(hipMemcpy(tmp.data(), segs[s].data + off, n * sizeof(T), hipMemcpyDeviceToHost), "D2H");
    out = std::copy(tmp.begin(), tmp.end(), out);
    i += n;
  }
  return out;
}
template <typename It, typename OutIt>
  requires requires(It i) { i.vec(); }
OutIt copy(It first, It last, OutIt out) {
  using T = typename std::iterator_traits<It>::value_type;
  return copy<T, OutIt>(first, last, out);
}

// Host -> device copy of the contiguous host range [first, last) into a
// distributed vector from position `out` on.
template <std::contiguous_iterator InIt, typename It>
  requires requires(It i) { i.vec(); i.pos(); }
It copy(InIt first, InIt last, It out) {
  using T = typename std::iterator_traits<It>::value_type;
  auto segs = out.vec()->segments();
  const std::size_t seg = out.vec()->segment_size();
  const std::size_t total = static_cast<std::size_t>(last - first);
  const T *src = std::to_address(first);
  std::size_t done = 0;
  for (std::size_t i = out.pos(); done < total;) {
    const std::size_t s = i / seg, off = i % seg;
    const std::size_t n = std::min(segs[s].size - off, total - done);
    dr::drx::hip_check(hipMemcpy(segs[s].data + off, src + done, n * sizeof(T),
                                 hipMemcpyHostToDevice), "H2D");
    done += n;
    i += n;
  }
  return out + static_cast<std::ptrdiff_t>(total);
}

template <typename T> class duplicated_vector {
public:
  using value_type = T;
  explicit duplicated_vector(std::size_t count = 0) : size_(count) {
    for (std::size_t r = 0; r < nprocs(); r++) segments_.emplace_back(count, r);
  }
  std::size_t size() const noexcept { return size_; }
  device_vector<T> &local_vector(std::size_t rank) { return segments_[rank]; }
  const device_vector<T> &local_vector(std::size_t rank) const { return segments_[rank]; }

private:
  std::vector<device_vector<T>> segments_;
  std::size_t size_ = 0;
};

// ---------------------------------------------------------------------------
// Sparse matrix: CSR tiles in HBM
// ---------------------------------------------------------------------------
template <typename T, std::integral I = std::int64_t> class sparse_matrix {
public:
  using size_type = std::size_t;
  using key_type = dr::index<I>;
  using value_type = dr::matrix_entry<T, I>;
  using segment_type = dr::views::csr_matrix_view<T, I>;

  explicit sparse_matrix(key_type shape) : shape_(shape), partition_(default_partition()) { init_empty(); }
  sparse_matrix(key_type shape, double density)
      : shape_(shape), partition_(default_partition()) { init_random(density); }
  sparse_matrix(key_type shape, double density, const matrix_partition &partition)
      : shape_(shape), partition_(partition.clone()) { init_random(density); }
  sparse_matrix(key_type shape, const matrix_partition &partition)
      : shape_(shape), partition_(partition.clone()) { init_empty(); }

  size_type size() const noexcept { return total_nnz_; }
  key_type shape() const noexcept { return shape_; }
  key_type grid_shape() const noexcept { return grid_; }
  key_type tile_shape() const noexcept { return tile_; }

  // The tile at (i, j): device pointers, local 0-based rowptr with tm + 1
  // entries, columns local to the tile's column range.
  segment_type tile(key_type t) const {
    const std::size_t idx = t[0] * grid_[1] + t[1];
    const auto &tl = tiles_[idx];
    return segment_type(tl.vals.data(), tl.rowptr.data(), tl.colind.data(),
                        key_type(tl.tm, tl.tn), tl.nnz, tl.rank);
  }
  std::vector<segment_type> tiles() const {
    std::vector<segment_type> out;
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) out.push_back(tile({i, j}));
    return out;
  }

  // Replace tile t by the host CSR `v` (shape of the tile).
  template <typename V> void copy_tile(key_type t, const V &v) {
    const std::size_t idx = t[0] * grid_[1] + t[1];
    auto &tl = tiles_[idx];
    total_nnz_ -= tl.nnz;
    tl.nnz = v.size();
    total_nnz_ += tl.nnz;
    tl.vals = dr::drx::device_buffer<T>(std::max<std::size_t>(tl.nnz, 1));
    tl.colind = dr::drx::device_buffer<I>(std::max<std::size_t>(tl.nnz, 1));
    tl.rowptr = dr::drx::device_buffer<I>(tl.tm + 1);
    tl.vals.upload(v.values_data(), tl.nnz);
    tl.colind.upload(v.colind_data(), tl.nnz);
    tl.rowptr.upload(v.rowptr_data(), tl.tm + 1);
    tl.plan = dr::drx::plan();
    snapshot_.reset();
  }

  // The kept plan of tile t (built on first use on stream q, then reused by
  // every product with the tile until copy_tile replaces it). The
  // reference's tiles persist across calls likewise
  // (sparse_matrix.hpp:193-207); only oneMKL's handle is rebuilt per call
  // (local_gemv.hpp:63-72), which this plan replaces.
  drx_plan tile_plan(key_type t, hipStream_t q) const {
    const std::size_t idx = t[0] * grid_[1] + t[1];
    const auto &tl = tiles_[idx];
    if (tl.plan.get() == nullptr && tl.tm > 0) {
      tl.plan = dr::drx::plan(tl.rowptr.data(), 0, static_cast<std::int64_t>(tl.tm),
                              static_cast<std::int64_t>(tl.tn),
                              static_cast<std::int64_t>(tl.nnz), true, q);
      tl.plan.auto_panels(dr::drx::dtype_code<T>(), tl.colind.data(), tl.vals.data(), q);
    }
    return tl.plan.get();
  }

  // Entries in tile order with global indices (downloads the tiles once).
  auto begin() const { return entries().begin(); }
  auto end() const { return entries().end(); }

private:
  struct tile_store {
    std::size_t tm = 0, tn = 0, nnz = 0, rank = 0, row0 = 0, col0 = 0;
    dr::drx::device_buffer<T> vals;
    dr::drx::device_buffer<I> rowptr, colind;
    mutable dr::drx::plan plan;  // built by tile_plan on first use
  };

  static std::unique_ptr<matrix_partition> default_partition() {
    return std::make_unique<block_cyclic>(row_cyclic());
  }

  void layout() {
    grid_ = key_type(partition_->grid_shape(dr::index<>(shape_[0], shape_[1])));
    tile_ = key_type(partition_->tile_shape(dr::index<>(shape_[0], shape_[1])));
    tiles_.clear();
    tiles_.resize(grid_[0] * grid_[1]);
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) {
        auto &tl = tiles_[i * grid_[1] + j];
        tl.tm = std::min<std::size_t>(tile_[0], shape_[0] - i * tile_[0]);
        tl.tn = std::min<std::size_t>(tile_[1], shape_[1] - j * tile_[1]);
        tl.row0 = i * tile_[0];
        tl.col0 = j * tile_[1];
        tl.rank = partition_->tile_rank(dr::index<>(shape_[0], shape_[1]), dr::index<>(i, j));
      }
  }

  void init_empty() {
    layout();
    for (auto &tl : tiles_) {
      std::vector<I> zeros(tl.tm + 1, 0);
      tl.rowptr = dr::drx::device_buffer<I>(tl.tm + 1);
      tl.rowptr.upload(zeros.data(), zeros.size());
      tl.vals = dr::drx::device_buffer<T>(1);
      tl.colind = dr::drx::device_buffer<I>(1);
    }
  }

  // sparse_matrix.hpp:348-403: tile (i, j) is generate_random_csr of the
  // tile's shape with seed i * grid[1] + j.
  void init_random(double density) {
    layout();
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) {
        auto &tl = tiles_[i * grid_[1] + j];
        auto csr = dr::generate_random_csr<T, I>(key_type(tl.tm, tl.tn), density,
                                                 static_cast<unsigned>(i * grid_[1] + j));
        copy_tile(key_type(i, j), csr);
        dr::destroy_csr_matrix_view(csr);
      }
  }

  const std::vector<value_type> &entries() const {
    if (!snapshot_) {
      auto out = std::make_shared<std::vector<value_type>>();
      for (const auto &tl : tiles_) {
        std::vector<T> v(tl.nnz);
        std::vector<I> c(tl.nnz), rp(tl.tm + 1);
        tl.vals.download(v.data(), tl.nnz);
        tl.colind.download(c.data(), tl.nnz);
        tl.rowptr.download(rp.data(), tl.tm + 1);
        for (std::size_t r = 0; r < tl.tm; r++)
          for (I q = rp[r]; q < rp[r + 1]; q++)
            out->emplace_back(key_type(tl.row0 + r, tl.col0 + c[q]), v[q]);
      }
      snapshot_ = out;
    }
    return *snapshot_;
  }

  key_type shape_, grid_, tile_;
  std::unique_ptr<matrix_partition> partition_;
  std::vector<tile_store> tiles_;
  std::size_t total_nnz_ = 0;
  mutable std::shared_ptr<std::vector<value_type>> snapshot_;
};

// Cut a host CSR into the partition's tiles (columns made tile-local).
template <typename T, typename I>
sparse_matrix<T, I> create_distributed(dr::views::csr_matrix_view<T, I> m,
                                       const matrix_partition &partition) {
  sparse_matrix<T, I> a(m.shape(), partition);
  const auto g = a.grid_shape(), ts = a.tile_shape();
  const I *rp = m.rowptr_data();
  const I *ci = m.colind_data();
  const T *va = m.values_data();
  for (std::size_t i = 0; i < g[0]; i++)
    for (std::size_t j = 0; j < g[1]; j++) {
      const std::size_t r0 = i * ts[0], r1 = std::min<std::size_t>(r0 + ts[0], m.shape()[0]);
      const std::size_t c0 = j * ts[1], c1 = std::min<std::size_t>(c0 + ts[1], m.shape()[1]);
      std::vector<T> v;
      std::vector<I> c, p{0};
      for (std::size_t r = r0; r < r1; r++) {
        for (I q = rp[r]; q < rp[r + 1]; q++)
          if (static_cast<std::size_t>(ci[q]) >= c0 && static_cast<std::size_t>(ci[q]) < c1) {
            v.push_back(va[q]);
            c.push_back(static_cast<I>(ci[q] - c0));
          }
        p.push_back(static_cast<I>(v.size()));
      }
      a.copy_tile({i, j}, dr::views::csr_matrix_view<T, I>(
                              v.data(), p.data(), c.data(), dr::index<I>(r1 - r0, c1 - c0),
                              v.size(), 0));
    }
  return a;
}

template <typename T, typename I>
sparse_matrix<T, I> create_distributed(dr::views::csr_matrix_view<T, I> m) {
  return create_distributed(m, row_cyclic());
}

// Read a Matrix Market file and distribute it
// (include/dr/sp/util/matrix_io.hpp:68-84).
template <typename T, typename I = std::size_t>
sparse_matrix<T, I> mmread(std::string file_path, const matrix_partition &partition,
                           bool one_indexed = true) {
  auto local = dr::read_csr<T, I>(file_path, one_indexed);
  auto a = create_distributed(local, partition);
  dr::__detail::destroy_csr_matrix_view(local, std::allocator<T>{});
  return a;
}

template <typename T, typename I = std::size_t>
sparse_matrix<T, I> mmread(std::string file_path, bool one_indexed = true) {
  return mmread<T, I>(file_path, row_cyclic(), one_indexed);
}

// ---------------------------------------------------------------------------
// gemv
// ---------------------------------------------------------------------------
namespace __detail {
// c[0 .. tm) += A_tile b on `q` (local_gemv.hpp:83-98 semantics).
template <typename T, typename I>
void local_gemv(hipStream_t q, dr::views::csr_matrix_view<T, I> a, const T *b, T *c) {
  if (a.shape()[0] == 0) return;
  dr::drx::check(drx_spmv_csr(dr::drx::itype_code<I>(), a.rowptr_data(), 0,
                              static_cast<std::int64_t>(a.shape()[0]),
                              static_cast<std::int64_t>(a.shape()[1]),
                              static_cast<std::int64_t>(a.size()), 1, a.colind_data(),
                              dr::drx::dtype_code<T>(), a.values_data(),
                              dr::drx::dtype_code<T>(), b,
                              static_cast<std::int64_t>(a.shape()[1]),
                              dr::drx::dtype_code<T>(), c,
                              static_cast<std::int64_t>(a.shape()[0]), 1, 1.0, 1.0, q),
                 "drx_spmv_csr");
}

// c[0 .. tm) += A_tile b through the tile's kept plan.
template <typename T, typename I>
void local_gemv(hipStream_t q, dr::views::csr_matrix_view<T, I> a, drx_plan plan,
                const T *b, T *c) {
  if (a.shape()[0] == 0) return;
  dr::drx::check(drx_spmv(plan, dr::drx::dtype_code<T>(), a.colind_data(),
                          a.values_data(),
                          dr::drx::dtype_code<T>(), b,
                          static_cast<std::int64_t>(a.shape()[1]),
                          dr::drx::dtype_code<T>(), c,
                          static_cast<std::int64_t>(a.shape()[0]), 1, 1.0, 1.0, q),
                 "drx_spmv");
}
}  // namespace __detail

// c += A b with row tiles ({P, 1} grid): b is replicated onto every device
// (scratch), then each tile's product accumulates into c's segment on the
// tile's device stream; returns after all streams finish.
template <typename T, typename I>
void gemv(distributed_vector<T> &c, sparse_matrix<T, I> &a, distributed_vector<T> &b,
          duplicated_vector<T> &scratch) {
  assert(a.shape()[0] == c.size());
  assert(a.shape()[1] == b.size());
  assert(a.grid_shape()[0] == c.segments().size());
  assert(a.grid_shape()[1] == 1);
  for (std::size_t r = 0; r < nprocs(); r++) {
    T *dst = scratch.local_vector(r).data();
    std::size_t off = 0;
    for (auto &s : b.segments()) {
      dr::drx::hip_check(hipMemcpyAsync(dst + off, s.data, s.size * sizeof(T),
                                        hipMemcpyDeviceToDevice, __detail::queue(r)),
                         "duplicate b");
      off += s.size;
    }
  }
  auto csegs = c.segments();
  for (std::size_t i = 0; i < a.grid_shape()[0]; i++) {
    auto t = a.tile(dr::index<I>(i, 0));
    assert(csegs[i].size == t.shape()[0]);
    hipStream_t q = __detail::queue(t.rank());
    __detail::local_gemv(q, t, a.tile_plan(dr::index<I>(i, 0), q),
                         scratch.local_vector(t.rank()).data(), csegs[i].data);
  }
  __detail::wait_all();
}

template <typename T, typename I>
void gemv(distributed_vector<T> &c, sparse_matrix<T, I> &a, distributed_vector<T> &b) {
  duplicated_vector<T> b_duplicated(b.size());
  gemv(c, a, b, b_duplicated);
}

// flat_gemv (gemv.hpp:19-76): the same product with row tiles ({P, 1} grid),
// b copied onto every device for this call. The reference walks each tile's
// entries with one atomic add per entry; here each tile runs the CSR kernel
// (accumulating, as gemv does).
template <typename T, typename I>
void flat_gemv(distributed_vector<T> &c, sparse_matrix<T, I> &a, distributed_vector<T> &b) {
  assert(a.grid_shape()[1] == 1);
  duplicated_vector<T> local_b(b.size());
  gemv(c, a, b, local_b);
}

namespace __detail {
// The tiles of a 2-D grid ({g0, g1}): row i takes its tiles in the rotation
// k = (k_ + i) % g1 (gemv.hpp:133-136), each on its device's stream. The
// reference's per-tile kernels add into c's segment with atomics, so the
// tiles of a row may overlap; here a row's tiles are chained with events,
// so one accumulating launch at a time writes a segment.
template <typename T, typename I, typename BPtr>
void gemv_tiles_2d(distributed_vector<T> &c, sparse_matrix<T, I> &a, BPtr &&b_of,
                   bool copy_segments) {
  const auto g = a.grid_shape();
  auto csegs = c.segments();
  std::vector<hipEvent_t> events;
  for (std::size_t i = 0; i < g[0]; i++) {
    hipEvent_t prev = nullptr;
    for (std::size_t k_ = 0; k_ < g[1]; k_++) {
      const std::size_t k = (k_ + i) % g[1];
      auto t = a.tile(dr::index<I>(i, k));
      hipStream_t q = queue(t.rank());
      if (prev) dr::drx::hip_check(hipStreamWaitEvent(q, prev, 0), "hipStreamWaitEvent");
      const T *bk = b_of(i, k, q, copy_segments);
      assert(csegs[i].size == t.shape()[0]);
      local_gemv(q, t, a.tile_plan(dr::index<I>(i, k), q), bk, csegs[i].data);
      hipEvent_t e = nullptr;
      dr::drx::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      dr::drx::hip_check(hipEventRecord(e, q), "hipEventRecord");
      events.push_back(e);
      prev = e;
    }
  }
  wait_all();
  for (auto e : events) (void)hipEventDestroy(e);
}
}  // namespace __detail

// gemv_square (gemv.hpp:126-156): c += A b over a 2-D tile grid; b's
// segment k is the right-hand side of tile column k (segments match the
// tile width).
template <typename T, typename I>
void gemv_square(distributed_vector<T> &c, sparse_matrix<T, I> &a, distributed_vector<T> &b) {
  assert(a.shape()[0] == c.size());
  assert(a.shape()[1] == b.size());
  assert(a.grid_shape()[0] == c.segments().size());
  assert(a.grid_shape()[1] == b.segments().size());
  auto bsegs = b.segments();
  __detail::gemv_tiles_2d(
      c, a, [&](std::size_t, std::size_t k, hipStream_t, bool) -> const T * {
        return bsegs[k].data;
      }, false);
}

// gemv_square_copy (gemv.hpp:158-207): as gemv_square, with each tile row's
// b segments first copied into a full-length b on that row's device.
template <typename T, typename I>
void gemv_square_copy(distributed_vector<T> &c, sparse_matrix<T, I> &a,
                      distributed_vector<T> &b) {
  assert(a.shape()[0] == c.size());
  assert(a.shape()[1] == b.size());
  assert(a.grid_shape()[0] == c.segments().size());
  assert(a.grid_shape()[1] == b.segments().size());
  auto bsegs = b.segments();
  const std::size_t tw = a.tile_shape()[1];
  std::vector<device_vector<T>> local_b;
  local_b.reserve(a.grid_shape()[0]);
  for (std::size_t i = 0; i < a.grid_shape()[0]; i++)
    local_b.emplace_back(b.size(), a.tile(dr::index<I>(i, 0)).rank());
  __detail::gemv_tiles_2d(
      c, a,
      [&](std::size_t i, std::size_t k, hipStream_t q, bool) -> const T * {
        T *dst = local_b[i].data() + k * tw;
        dr::drx::hip_check(hipMemcpyAsync(dst, bsegs[k].data, bsegs[k].size * sizeof(T),
                                          hipMemcpyDeviceToDevice, q),
                           "copy b segment");
        return dst;
      },
      true);
}

// ---------------------------------------------------------------------------
// Dense matrices and their product
// ---------------------------------------------------------------------------

// The three partitions of c = a b over the processor grid factor(nprocs()):
// c tiled by that grid, a and b sharing a k split sized to cover every rank.
inline std::vector<block_cyclic> partition_matmul(std::size_t m, std::size_t n, std::size_t k) {
  const dr::index<> c_grid = detail::factor(nprocs());
  const std::size_t k_block = (m * k >= k * n) ? (nprocs() + c_grid[0] - 1) / c_grid[0]
                                               : (nprocs() + c_grid[1] - 1) / c_grid[1];
  return {block_cyclic({tile::div, tile::div}, {c_grid[0], k_block}),
          block_cyclic({tile::div, tile::div}, {k_block, c_grid[1]}),
          block_cyclic({tile::div, tile::div}, {c_grid[0], c_grid[1]})};
}

// A row-major window of device memory: shape, leading dimension, owner.
template <typename T> class dense_matrix_view {
public:
  using key_type = dr::index<>;
  dense_matrix_view(T *data, key_type shape, std::size_t ld, std::size_t rank)
      : data_(data), shape_(shape), ld_(ld), rank_(rank) {}
  T *data() const noexcept { return data_; }
  key_type shape() const noexcept { return shape_; }
  std::size_t size() const noexcept { return shape_[0] * shape_[1]; }
  std::size_t ld() const noexcept { return ld_; }
  std::size_t rank() const noexcept { return rank_; }

private:
  T *data_;
  key_type shape_;
  std::size_t ld_;
  std::size_t rank_;
};

// Dense matrix cut into the partition's tiles, each a row-major block in HBM
// placed on the tile's rank. Tiles start zeroed. Their rows are padded to a
// multiple of 32 elements so that every row starts 16-byte aligned for the
// matrix-core loads; ld() of a tile view reports the padded length.
template <typename T> class distributed_dense_matrix {
public:
  using key_type = dr::index<>;
  using size_type = std::size_t;

  class element_ref {
  public:
    element_ref(T *p) : p_(p) {}
    operator T() const {
      T v;
      dr::drx::hip_check(hipMemcpy(&v, p_, sizeof(T), hipMemcpyDeviceToHost), "read element");
      return v;
    }
    element_ref &operator=(T v) {
      dr::drx::hip_check(hipMemcpy(p_, &v, sizeof(T), hipMemcpyHostToDevice), "write element");
      return *this;
    }

  private:
    T *p_;
  };

  explicit distributed_dense_matrix(key_type shape)
      : distributed_dense_matrix(shape, block_cyclic()) {}
  distributed_dense_matrix(key_type shape, const matrix_partition &partition)
      : shape_(shape), partition_(partition.clone()) {
    grid_ = partition_->grid_shape(shape_);
    tile_ = partition_->tile_shape(shape_);
    ld_ = (tile_[1] + 31) / 32 * 32;
    tiles_.reserve(grid_[0] * grid_[1]);
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) {
        tiles_.emplace_back(tile_[0] * ld_, partition_->tile_rank(shape_, key_type(i, j)));
        if (tiles_.back().size())
          dr::drx::hip_check(hipMemset(tiles_.back().data(), 0, tiles_.back().size() * sizeof(T)),
                             "zero tile");
      }
  }

  size_type size() const noexcept { return shape_[0] * shape_[1]; }
  key_type shape() const noexcept { return shape_; }
  key_type tile_shape() const noexcept { return tile_; }
  key_type grid_shape() const noexcept { return grid_; }

  dense_matrix_view<T> tile(key_type t) const {
    const std::size_t tm = std::min(tile_[0], shape_[0] - t[0] * tile_[0]);
    const std::size_t tn = std::min(tile_[1], shape_[1] - t[1] * tile_[1]);
    const auto &tv = tiles_[t[0] * grid_[1] + t[1]];
    return dense_matrix_view<T>(tv.data(), key_type(tm, tn), ld_, tv.rank());
  }

  std::vector<dense_matrix_view<T>> tiles() const {
    std::vector<dense_matrix_view<T>> v;
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) v.push_back(tile(key_type(i, j)));
    return v;
  }

  // Element access through host copies (for setting up and checking).
  element_ref operator[](key_type idx) {
    const auto &tv = tiles_[idx[0] / tile_[0] * grid_[1] + idx[1] / tile_[1]];
    return element_ref(tv.data() + (idx[0] % tile_[0]) * ld_ + idx[1] % tile_[1]);
  }

  // Whole-matrix copies from and to a dense row-major host array.
  void copy_from_host(const T *host) {
    for_tiles([&](dense_matrix_view<T> v, std::size_t r0, std::size_t c0) {
      dr::drx::hip_check(hipMemcpy2D(v.data(), v.ld() * sizeof(T), host + r0 * shape_[1] + c0,
                                     shape_[1] * sizeof(T), v.shape()[1] * sizeof(T),
                                     v.shape()[0], hipMemcpyHostToDevice),
                         "upload tile");
    });
  }
  void copy_to_host(T *host) const {
    for_tiles([&](dense_matrix_view<T> v, std::size_t r0, std::size_t c0) {
      dr::drx::hip_check(hipMemcpy2D(host + r0 * shape_[1] + c0, shape_[1] * sizeof(T), v.data(),
                                     v.ld() * sizeof(T), v.shape()[1] * sizeof(T), v.shape()[0],
                                     hipMemcpyDeviceToHost),
                         "download tile");
    });
  }

private:
  template <typename F> void for_tiles(F &&f) const {
    for (std::size_t i = 0; i < grid_[0]; i++)
      for (std::size_t j = 0; j < grid_[1]; j++) {
        auto v = tile(key_type(i, j));
        if (v.size()) f(v, i * tile_[0], j * tile_[1]);
      }
  }

  key_type shape_, grid_, tile_;
  std::size_t ld_ = 0;
  std::unique_ptr<matrix_partition> partition_;
  std::vector<device_vector<T>> tiles_;
};

namespace __detail {
template <typename T> constexpr int dense_vtype() {
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>,
                "the dense product takes float or double");
  return std::is_same_v<T, float> ? DRX_F32 : DRX_F64;
}

// c += a b on one tile, asynchronously on q (drx_gemm, alpha = beta = 1).
template <typename T>
void local_gemm(hipStream_t q, dense_matrix_view<T> a, dense_matrix_view<T> b,
                dense_matrix_view<T> c) {
  assert(c.shape()[0] == a.shape()[0]);
  assert(c.shape()[1] == b.shape()[1]);
  assert(a.shape()[1] == b.shape()[0]);
  dr::drx::check(drx_gemm(dense_vtype<T>(), c.shape()[0], c.shape()[1], a.shape()[1], 1.0,
                          a.data(), std::max<std::size_t>(a.ld(), 1), b.data(),
                          std::max<std::size_t>(b.ld(), 1), 1.0, c.data(),
                          std::max<std::size_t>(c.ld(), 1), q),
                 "drx_gemm");
}
}  // namespace __detail

// c += a b. Each c tile's products run in order on its rank's stream, k
// starting at (i + j) mod grid_k as the reference staggers them. The
// reference's buffered forms copy remote a and b tiles into local buffers
// first; on one GPU every tile is already in local HBM, so all four forms
// issue the same tile products.
template <typename T>
void gemm_inplace(distributed_dense_matrix<T> &a, distributed_dense_matrix<T> &b,
                  distributed_dense_matrix<T> &c) {
  assert(c.shape()[0] == a.shape()[0]);
  assert(c.shape()[1] == b.shape()[1]);
  assert(a.shape()[1] == b.shape()[0]);
  assert(c.grid_shape()[0] == a.grid_shape()[0]);
  assert(c.grid_shape()[1] == b.grid_shape()[1]);
  assert(a.grid_shape()[1] == b.grid_shape()[0]);
  const std::size_t gk = a.grid_shape()[1];
  for (std::size_t i = 0; i < c.grid_shape()[0]; i++)
    for (std::size_t j = 0; j < c.grid_shape()[1]; j++) {
      auto ct = c.tile(dr::index<>(i, j));
      hipStream_t q = __detail::queue(ct.rank());
      for (std::size_t k_ = 0; k_ < gk; k_++) {
        const std::size_t k = (k_ + i + j) % gk;
        __detail::local_gemm(q, a.tile(dr::index<>(i, k)), b.tile(dr::index<>(k, j)), ct);
      }
    }
  __detail::wait_all();
}

template <typename T>
void gemm_buffered(distributed_dense_matrix<T> &a, distributed_dense_matrix<T> &b,
                   distributed_dense_matrix<T> &c) {
  gemm_inplace(a, b, c);
}

template <typename T>
void gemm_buffered_async(distributed_dense_matrix<T> &a, distributed_dense_matrix<T> &b,
                         distributed_dense_matrix<T> &c) {
  gemm_inplace(a, b, c);
}

template <typename T>
void gemm(distributed_dense_matrix<T> &a, distributed_dense_matrix<T> &b,
          distributed_dense_matrix<T> &c) {
  gemm_buffered(a, b, c);
}

}  // namespace dr::sp
