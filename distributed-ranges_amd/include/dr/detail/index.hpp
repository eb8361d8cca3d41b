// dr/detail/index.hpp -- 2-D index and matrix entry types of the host
// interface (the reference's dr::index / dr::matrix_entry, include/dr/detail/
// index.hpp and include/dr/detail/matrix_entry.hpp, same member names).
#pragma once

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <tuple>
#include <utility>

namespace dr {

template <std::integral I = std::size_t> struct index {
  using index_type = I;
  I first = 0;
  I second = 0;

  constexpr index() noexcept = default;
  template <std::integral A, std::integral B>
  constexpr index(A a, B b) noexcept
      : first(static_cast<I>(a)), second(static_cast<I>(b)) {}
  template <std::integral J>
  constexpr index(index<J> o) noexcept
      : first(static_cast<I>(o.first)), second(static_cast<I>(o.second)) {}

  constexpr I &operator[](std::size_t d) noexcept { return d == 0 ? first : second; }
  constexpr const I &operator[](std::size_t d) const noexcept {
    return d == 0 ? first : second;
  }
  template <std::size_t N> constexpr I get() const noexcept {
    return N == 0 ? first : second;
  }
  constexpr bool operator==(const index &) const noexcept = default;
  constexpr auto operator<=>(const index &) const noexcept = default;
};

template <typename T, std::integral I = std::size_t> struct matrix_entry {
  using index_type = I;
  using value_type = T;
  dr::index<I> idx;
  T val;

  matrix_entry() = default;
  matrix_entry(dr::index<I> i, T v) : idx(i), val(v) {}
  dr::index<I> index() const noexcept { return idx; }
  T value() const noexcept { return val; }
  template <std::size_t N> decltype(auto) get() const noexcept {
    if constexpr (N == 0)
      return idx;
    else
      return val;
  }
  bool operator<(const matrix_entry &o) const noexcept { return idx < o.idx; }
};

}  // namespace dr

template <std::integral I>
struct std::tuple_size<dr::index<I>> : std::integral_constant<std::size_t, 2> {};
template <std::size_t N, std::integral I>
struct std::tuple_element<N, dr::index<I>> {
  using type = I;
};
template <typename T, std::integral I>
struct std::tuple_size<dr::matrix_entry<T, I>>
    : std::integral_constant<std::size_t, 2> {};
template <typename T, std::integral I>
struct std::tuple_element<0, dr::matrix_entry<T, I>> {
  using type = dr::index<I>;
};
template <typename T, std::integral I>
struct std::tuple_element<1, dr::matrix_entry<T, I>> {
  using type = T;
};
