// dr/detail/matrix_io.hpp -- COO -> CSR conversion of sorted (row, col)
// tuples and the Matrix Market reader, the interface of
// include/dr/detail/matrix_io.hpp:26-72 (dr::__detail::convert_to_csr),
// :230-239 (dr::__detail::destroy_csr_matrix_view) and :243-252
// (dr::read_csr). The file is parsed by the library's parallel host reader
// (drx_mm_read_csr).
#pragma once

#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "drx_spmv.h"
#include "dr/views/csr_matrix_view.hpp"

namespace dr::__detail {

// `tuples`: ((row, col), value) entries sorted by (row, col); the arrays are
// allocated with `allocator` (rebound for the index type).
template <typename Tuples, typename Allocator>
auto convert_to_csr(Tuples &&tuples, dr::index<> shape, std::size_t nnz,
                    Allocator &&allocator) {
  using entry = std::remove_cvref_t<decltype(*std::begin(tuples))>;
  using T = std::remove_cvref_t<decltype(std::declval<entry>().second)>;
  using I = std::remove_cvref_t<decltype(std::declval<entry>().first.first)>;
  typename std::allocator_traits<std::remove_cvref_t<Allocator>>::template rebind_alloc<I>
      ialloc(allocator);
  T *values = allocator.allocate(nnz);
  I *rowptr = ialloc.allocate(shape[0] + 1);
  I *colind = ialloc.allocate(nnz);
  std::size_t row = 0, q = 0;
  rowptr[0] = 0;
  for (auto &&[ij, v] : tuples) {
    values[q] = v;
    colind[q] = ij.second;
    for (; row < static_cast<std::size_t>(ij.first); row++) rowptr[row + 1] = static_cast<I>(q);
    q++;
  }
  for (; row < shape[0]; row++) rowptr[row + 1] = static_cast<I>(nnz);
  return dr::views::csr_matrix_view<T, I>(values, rowptr, colind,
                                          dr::index<I>(shape[0], shape[1]), nnz, 0);
}

// Release a view whose arrays were allocated with `alloc` (rebound for the
// index type), as read_csr's are with std::allocator<T>.
template <typename T, typename I, typename Allocator>
void destroy_csr_matrix_view(dr::views::csr_matrix_view<T, I> view, Allocator &&alloc) {
  alloc.deallocate(view.values_data(), view.size());
  typename std::allocator_traits<std::remove_cvref_t<Allocator>>::template rebind_alloc<I>
      ialloc(alloc);
  ialloc.deallocate(view.colind_data(), view.size());
  ialloc.deallocate(view.rowptr_data(), view.shape()[0] + 1);
}

}  // namespace dr::__detail

namespace dr {

// Read a coordinate Matrix Market file into a host CSR view (0-based rowptr
// with sentinel, rows sorted by column); the arrays come from
// std::allocator: release them with
// dr::__detail::destroy_csr_matrix_view(view, std::allocator<T>{}). Throws
// std::runtime_error with the reader's message on failure.
template <typename T, typename I = std::size_t>
auto read_csr(std::string file_path, bool one_indexed = true) {
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>,
                "read_csr reads float or double values");
  static_assert(std::is_integral_v<I> && (sizeof(I) == 4 || sizeof(I) == 8),
                "read_csr reads 32- or 64-bit indices");
  constexpr int vt = std::is_same_v<T, float> ? DRX_F32 : DRX_F64;
  constexpr int it = sizeof(I) == 4 ? DRX_I32 : DRX_I64;
  std::int64_t m = 0, n = 0, nnz = 0;
  void *rp = nullptr, *ci = nullptr, *va = nullptr;
  if (drx_mm_read_csr(file_path.c_str(), one_indexed ? 1 : 0, it, vt, 0, &m, &n, &nnz, &rp,
                      &ci, &va) != DRX_OK)
    throw std::runtime_error(drx_last_error());
  std::allocator<T> alloc;
  std::allocator<I> ialloc;
  T *values = alloc.allocate(static_cast<std::size_t>(nnz));
  I *rowptr = ialloc.allocate(static_cast<std::size_t>(m) + 1);
  I *colind = ialloc.allocate(static_cast<std::size_t>(nnz));
  std::memcpy(values, va, static_cast<std::size_t>(nnz) * sizeof(T));
  std::memcpy(rowptr, rp, (static_cast<std::size_t>(m) + 1) * sizeof(I));
  std::memcpy(colind, ci, static_cast<std::size_t>(nnz) * sizeof(I));
  drx_host_free(rp);
  drx_host_free(ci);
  drx_host_free(va);
  return dr::views::csr_matrix_view<T, I>(values, rowptr, colind,
                                          dr::index<I>(static_cast<I>(m), static_cast<I>(n)),
                                          static_cast<std::size_t>(nnz), 0);
}

}  // namespace dr
