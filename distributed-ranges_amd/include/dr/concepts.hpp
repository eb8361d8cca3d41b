// dr/concepts.hpp -- the range concepts the gemv path's containers satisfy.
//
//   dr::ranges::segments(r)   the segments of a distributed range (its
//                             member segments()), as the reference's
//                             customization point (include/dr/detail/
//                             ranges.hpp) resolves it for these containers
//   dr::distributed_range<R>  a forward range with segments
//                             (include/dr/concepts/concepts.hpp:20-21)
#pragma once

#include <ranges>

namespace dr {

namespace ranges {

struct segments_fn {
  template <typename R>
    requires requires(R &&r) { r.segments(); }
  decltype(auto) operator()(R &&r) const {
    return r.segments();
  }
};
inline constexpr segments_fn segments{};

}  // namespace ranges

template <typename R>
concept distributed_range =
    std::ranges::forward_range<R> && requires(R &r) { dr::ranges::segments(r); };

}  // namespace dr
