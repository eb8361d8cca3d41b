// dr/detail/generate_csr.hpp -- host generators of test/benchmark matrices.
//
// Same entry points and draws as the reference's generators
// (include/dr/detail/generate_csr.hpp:40-103 generate_random_csr,
// :108-153 generate_band_csr): a std::mt19937 seeded with `seed`, row and
// column from std::uniform_int_distribution, values from
// std::uniform_{int,real}_distribution(0, 1), duplicates redrawn, entries
// sorted by (row, column). The draw order has to be the same for the same
// standard library to produce the same matrix as the reference's tests. The
// memory is owned by the caller: release it with destroy_csr_matrix_view.
#pragma once

#include <algorithm>
#include <cassert>
#include <concepts>
#include <cstdint>
#include <random>
#include <unordered_set>
#include <utility>
#include <vector>

#include "dr/views/csr_matrix_view.hpp"

namespace dr {

namespace __detail {
template <typename T> struct unit_draw {
  using type = std::uniform_int_distribution<T>;
};
template <std::floating_point T> struct unit_draw<T> {
  using type = std::uniform_real_distribution<T>;
};
struct cell_hash {
  std::size_t operator()(const std::pair<std::uint64_t, std::uint64_t> &c) const {
    return c.first * 31 + c.second;
  }
};
}  // namespace __detail

template <typename T = float, std::integral I = std::size_t>
auto generate_random_csr(dr::index<I> shape, double density = 0.01,
                         unsigned int seed = 0) {
  assert(density >= 0.0 && density < 1.0);
  const std::size_t m = shape[0], n = shape[1];
  const std::size_t nnz = density * m * n;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<I> pick_row(0, static_cast<I>(m - 1));
  std::uniform_int_distribution<I> pick_col(0, static_cast<I>(n - 1));
  typename __detail::unit_draw<T>::type pick_val(0, 1);
  std::unordered_set<std::pair<std::uint64_t, std::uint64_t>, __detail::cell_hash> seen;
  std::vector<std::pair<std::pair<I, I>, T>> cells;
  cells.reserve(nnz);
  while (seen.size() < nnz) {
    const I i = pick_row(gen);
    const I j = pick_col(gen);
    const std::pair<std::uint64_t, std::uint64_t> key(i, j);
    if (!seen.contains(key)) {
      const T v = pick_val(gen);
      seen.insert(key);
      cells.push_back({{i, j}, v});
    }
  }
  std::sort(cells.begin(), cells.end());
  T *values = new T[nnz];
  I *rowptr = new I[m + 1];
  I *colind = new I[nnz];
  std::size_t row = 0;
  rowptr[0] = 0;
  for (std::size_t q = 0; q < cells.size(); q++) {
    const auto &[ij, v] = cells[q];
    values[q] = v;
    colind[q] = ij.second;
    for (; row < static_cast<std::size_t>(ij.first); row++) rowptr[row + 1] = static_cast<I>(q);
  }
  for (; row < m; row++) rowptr[row + 1] = static_cast<I>(nnz);
  return dr::views::csr_matrix_view<T, I>(values, rowptr, colind, shape, nnz, 0);
}

template <typename T = float, std::integral I = std::size_t>
auto generate_band_csr(I size, std::size_t up_band = 3, std::size_t down_band = 3) {
  const std::size_t n = static_cast<std::size_t>(size);
  const std::size_t nnz = (1 + up_band + down_band) * n - up_band * (up_band + 1) / 2 -
                          down_band * (down_band + 1) / 2;
  T *values = new T[nnz];
  I *rowptr = new I[n + 1];
  I *colind = new I[nnz];
  std::size_t q = 0;
  rowptr[0] = 0;
  for (std::size_t i = 0; i < n; i++) {
    const std::size_t lo = i > down_band ? i - down_band : 0;
    const std::size_t hi = std::min(n - 1, i + up_band);
    for (std::size_t j = lo; j <= hi; j++, q++) {
      values[q] = 1;
      colind[q] = static_cast<I>(j);
    }
    rowptr[i + 1] = static_cast<I>(q);
  }
  return dr::views::csr_matrix_view<T, I>(values, rowptr, colind,
                                          dr::index<I>(size, size), nnz, 0);
}

template <typename T, typename I>
void destroy_csr_matrix_view(dr::views::csr_matrix_view<T, I> view) {
  delete[] view.values_data();
  delete[] view.rowptr_data();
  delete[] view.colind_data();
}

}  // namespace dr
