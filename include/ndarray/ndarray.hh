// Minimal ftk::ndarray<T> with the subset of hguo/ndarray's interface that the
// simplicial engine and its callers use (construction from extents, Fortran-order
// f(...) access, operator[], nd/dimf/shapef/nelem, fill, reshapef, pdata).
// Storage is contiguous with the first index fastest.
#pragma once
#include <algorithm>
#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <vector>

namespace ftk {

template <class T>
class ndarray {
 public:
  ndarray() = default;
  ndarray(std::initializer_list<size_t> dims) { reshapef(std::vector<size_t>(dims)); }
  explicit ndarray(const std::vector<size_t>& dims) { reshapef(dims); }

  void reshapef(const std::vector<size_t>& dims) {
    dims_ = dims;
    size_t n = 1;
    for (size_t d : dims_) n *= d;
    data_.assign(n, T(0));
  }
  size_t nd() const { return dims_.size(); }
  size_t dimf(size_t i) const { return dims_.at(i); }
  const std::vector<size_t>& shapef() const { return dims_; }
  size_t nelem() const { return data_.size(); }
  void fill(T v) { std::fill(data_.begin(), data_.end(), v); }
  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }
  const T* pdata() const { return data_.data(); }
  T& operator[](size_t i) { return data_[i]; }
  const T& operator[](size_t i) const { return data_[i]; }

  template <class... I>
  T& f(I... idx) {
    return data_[offset({static_cast<size_t>(idx)...})];
  }
  template <class... I>
  const T& f(I... idx) const {
    return data_[offset({static_cast<size_t>(idx)...})];
  }

 private:
  size_t offset(std::initializer_list<size_t> idx) const {
    size_t off = 0, mul = 1, j = 0;
    for (size_t i : idx) {
      off += i * mul;
      mul *= j < dims_.size() ? dims_[j] : 1;
      ++j;
    }
    return off;
  }
  std::vector<size_t> dims_;
  std::vector<T> data_;
};

}  // namespace ftk
