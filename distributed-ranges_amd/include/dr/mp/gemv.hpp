// dr/mp/gemv.hpp -- dr::mp::gemv, the four overloads of
// include/dr/mp/algorithms/matrix/gemv.hpp:16-71.
//
// Rooted forms compute res (on root) = A * b through the matrix's
// distribution; the all-ranks forms gather into a workspace on rank 0 and
// broadcast the result so every rank's res holds it. res is a host range of
// m (vector) or m * width (slim matrix, column-major) elements.
#pragma once

#include <algorithm>
#include <cassert>
#include <ranges>
#include <vector>

#include "dr/mp/broadcasted.hpp"
#include "dr/mp/sparse.hpp"

namespace dr::mp {

template <typename T, typename I, std::ranges::range C, typename Alloc,
          typename Backend, typename MatDistr>
  requires(vector_multiplicable<MatDistr>)
void gemv(int root, C &res, distributed_sparse_matrix<T, I, Backend, MatDistr> &a,
          broadcasted_vector<T, Alloc> b) {
  if (default_comm().rank() == static_cast<std::size_t>(root)) {
    assert(a.shape().first == std::ranges::size(res));
    assert(a.shape().second == b.size());
  }
  a.local_gemv_and_collect(root, res, b.broadcasted_data(), 1);
}

template <typename T, typename I, std::ranges::range C, typename Alloc,
          typename Backend, typename MatDistr>
  requires(vector_multiplicable<MatDistr>)
void gemv(int root, C &res, distributed_sparse_matrix<T, I, Backend, MatDistr> &a,
          broadcasted_slim_matrix<T, Alloc> b) {
  if (default_comm().rank() == static_cast<std::size_t>(root))
    assert(a.shape().first * b.width() == std::ranges::size(res));
  a.local_gemv_and_collect(root, res, b.broadcasted_data(), b.width());
}

template <typename T, typename I, std::ranges::range C, typename Alloc,
          typename Backend, typename MatDistr>
  requires(vector_multiplicable<MatDistr>)
void gemv(C &res, distributed_sparse_matrix<T, I, Backend, MatDistr> &a,
          broadcasted_vector<T, Alloc> b) {
  std::vector<T> workspace(std::ranges::size(res));
  gemv(0, workspace, a, b);
  default_comm().bcast(workspace.data(), sizeof(T) * workspace.size(), 0);
  std::copy(workspace.begin(), workspace.end(), std::ranges::begin(res));
}

template <typename T, typename I, std::ranges::range C, typename Alloc,
          typename Backend, typename MatDistr>
  requires(vector_multiplicable<MatDistr>)
void gemv(C &res, distributed_sparse_matrix<T, I, Backend, MatDistr> &a,
          broadcasted_slim_matrix<T, Alloc> b) {
  std::vector<T> workspace(std::ranges::size(res));
  gemv(0, workspace, a, b);
  default_comm().bcast(workspace.data(), sizeof(T) * workspace.size(), 0);
  std::copy(workspace.begin(), workspace.end(), std::ranges::begin(res));
}

}  // namespace dr::mp
