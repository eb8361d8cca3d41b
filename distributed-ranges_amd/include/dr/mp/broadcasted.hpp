// dr/mp/broadcasted.hpp -- the replicated right-hand side of dr::mp::gemv.
//
// broadcasted_vector (include/dr/mp/containers/broadcasted_vector.hpp:10-53)
// and broadcasted_slim_matrix (.../broadcasted_slim_matrix.hpp:10-83): every
// rank holds a full copy of x in device memory; the slim matrix holds `width`
// right-hand sides column-major ([width][height], element (i, j) at
// data[i + j * height]). broadcast_data is collective: root's host data is
// published to the team and each rank uploads it to its own device copy (the
// role of the reference's copy-to-USM + MPI_Bcast).
#pragma once

#include <cstddef>
#include <iterator>
#include <memory>
#include <ranges>

#include "dr/mp/runtime.hpp"

namespace dr::mp {

template <typename T, typename Allocator = __detail::allocator<T>>
class broadcasted_vector {
public:
  broadcasted_vector() = default;

  template <std::ranges::input_range R>
  void broadcast_data(std::size_t data_size, std::size_t root, R root_data,
                      communicator comm) {
    if (_data != nullptr) destroy_data();
    _data_size = data_size;
    _data = alloc.allocate(_data_size);
    const T *src = nullptr;
    if (comm.rank() == root) src = std::to_address(std::ranges::begin(root_data));
    comm.bcast(&src, sizeof(src), root);
    if (_data_size)
      dr::drx::hip_check(hipMemcpyAsync(_data, src, _data_size * sizeof(T),
                                        hipMemcpyHostToDevice, stream()),
                         "broadcast H2D");
    dr::drx::hip_check(hipStreamSynchronize(stream()), "sync");
    comm.barrier();
  }

  void destroy_data() {
    alloc.deallocate(_data, _data_size);
    _data_size = 0;
    _data = nullptr;
  }

  T *broadcasted_data() { return _data; }
  auto size() { return _data_size; }
  auto begin() const { return _data; }
  auto end() const { return begin() + _data_size; }

private:
  T *_data = nullptr;
  std::size_t _data_size = 0;
  Allocator alloc;
};

template <typename T, typename Allocator = __detail::allocator<T>>
class broadcasted_slim_matrix {
public:
  broadcasted_slim_matrix() = default;

  // width separate columns of `height` elements each.
  void broadcast_data(std::size_t height, std::size_t width, std::size_t root,
                      T **root_data, communicator comm) {
    reset_shape(height, width);
    T *const *cols = nullptr;
    if (comm.rank() == root) cols = root_data;
    comm.bcast(&cols, sizeof(cols), root);
    for (std::size_t i = 0; i < width; i++)
      if (height)
        dr::drx::hip_check(hipMemcpyAsync(_data + height * i, cols[i], height * sizeof(T),
                                          hipMemcpyHostToDevice, stream()),
                           "broadcast H2D");
    dr::drx::hip_check(hipStreamSynchronize(stream()), "sync");
    comm.barrier();
  }

  // A contiguous column-major range of height * width elements.
  template <std::ranges::input_range R>
  void broadcast_data(std::size_t height, std::size_t width, std::size_t root,
                      R root_data, communicator comm) {
    reset_shape(height, width);
    const T *src = nullptr;
    if (comm.rank() == root) src = std::to_address(std::ranges::begin(root_data));
    comm.bcast(&src, sizeof(src), root);
    if (_data_size)
      dr::drx::hip_check(hipMemcpyAsync(_data, src, _data_size * sizeof(T),
                                        hipMemcpyHostToDevice, stream()),
                         "broadcast H2D");
    dr::drx::hip_check(hipStreamSynchronize(stream()), "sync");
    comm.barrier();
  }

  void destroy_data() {
    alloc.deallocate(_data, _data_size);
    _data_size = 0;
    _data = nullptr;
  }

  T *operator[](std::size_t index) { return _data + _height * index; }
  T *broadcasted_data() { return _data; }
  auto width() { return _width; }
  auto height() { return _height; }

private:
  void reset_shape(std::size_t height, std::size_t width) {
    if (_data != nullptr) destroy_data();
    _data_size = height * width;
    _height = height;
    _width = width;
    _data = alloc.allocate(_data_size);
  }

  T *_data = nullptr;
  std::size_t _data_size = 0;
  std::size_t _width = 0;
  std::size_t _height = 0;
  Allocator alloc;
};

}  // namespace dr::mp
