// dr/views/csr_matrix_view.hpp -- non-owning CSR view: the input of
// dr::mp::distributed_sparse_matrix and the tile type of dr::sp::sparse_matrix.
//
// Interface of include/dr/views/csr_matrix_view.hpp:125-218 (values_data,
// rowptr_data, colind_data, shape, size, rank, row, iteration over
// (index, value) entries with an index offset); rowptr has shape[0] + 1
// entries. Iteration dereferences host memory only (views over device tiles
// are iterated through dr::sp::sparse_matrix, which downloads them).
#pragma once

#include <cstddef>
#include <iterator>

#include "dr/detail/index.hpp"

namespace dr::views {

template <typename T, typename I, typename TIter = T *, typename IIter = I *>
class csr_matrix_view {
public:
  using size_type = std::size_t;
  using difference_type = std::ptrdiff_t;
  using scalar_type = T;
  using index_type = I;
  using key_type = dr::index<I>;
  using value_type = dr::matrix_entry<T, I>;

  // Forward iterator over the entries in storage order.
  class iterator {
  public:
    using iterator_category = std::forward_iterator_tag;
    using value_type = dr::matrix_entry<T, I>;
    using difference_type = std::ptrdiff_t;
    iterator() = default;
    iterator(const csr_matrix_view *v, size_type idx) : v_(v), idx_(idx), row_(0) {
      advance_row();
    }
    value_type operator*() const {
      return value_type(key_type(row_ + v_->idx_offset_[0],
                                 v_->colind_[idx_] + v_->idx_offset_[1]),
                        v_->values_[idx_]);
    }
    iterator &operator++() {
      ++idx_;
      advance_row();
      return *this;
    }
    iterator operator++(int) {
      auto o = *this;
      ++*this;
      return o;
    }
    bool operator==(const iterator &o) const { return idx_ == o.idx_; }

  private:
    // The row holding entry idx_: advance past rows that end at or before it.
    void advance_row() {
      while (v_ && static_cast<size_type>(row_) + 1 < static_cast<size_type>(v_->shape_[0]) &&
             idx_ >= static_cast<size_type>(v_->rowptr_[row_ + 1]))
        ++row_;
    }
    const csr_matrix_view *v_ = nullptr;
    size_type idx_ = 0;
    I row_ = 0;
  };

  csr_matrix_view() = default;
  csr_matrix_view(TIter values, IIter rowptr, IIter colind, key_type shape,
                  size_type nnz, size_type rank, key_type idx_offset = key_type(0, 0))
      : values_(values), rowptr_(rowptr), colind_(colind), shape_(shape),
        nnz_(nnz), rank_(rank), idx_offset_(idx_offset) {}

  key_type shape() const noexcept { return shape_; }
  size_type size() const noexcept { return nnz_; }
  std::size_t rank() const noexcept { return rank_; }
  TIter values_data() const noexcept { return values_; }
  IIter rowptr_data() const noexcept { return rowptr_; }
  IIter colind_data() const noexcept { return colind_; }
  key_type idx_offset() const noexcept { return idx_offset_; }

  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, nnz_); }

private:
  TIter values_{};
  IIter rowptr_{};
  IIter colind_{};
  key_type shape_{};
  size_type nnz_ = 0;
  size_type rank_ = 0;
  key_type idx_offset_{};
};

}  // namespace dr::views
