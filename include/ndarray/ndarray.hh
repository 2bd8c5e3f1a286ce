// Minimal ftk::ndarray<T> with the subset of hguo/ndarray's interface that the
// simplicial engine and its callers use (construction from extents, Fortran-order
// f(...) access, operator[], nd/dimf/shapef/nelem, fill, reshapef, pdata), and the
// ftk::ndarray_group / ftk::stream<> read interface of execute_stream().
// Storage is contiguous with the first index fastest.
#pragma once
#include <algorithm>
#include <cstddef>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <typeinfo>
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

// Named arrays of one timestep (get_ref<T>(name), has(name)).
class ndarray_group {
 public:
  template <class T>
  void set(const std::string& name, ndarray<T> a) {
    arrays_[name] = Entry{std::type_index(typeid(T)), std::make_shared<ndarray<T>>(std::move(a))};
  }
  bool has(const std::string& name) const { return arrays_.count(name) != 0; }
  template <class T>
  const ndarray<T>& get_ref(const std::string& name) const {
    const Entry& e = arrays_.at(name);
    if (e.type != std::type_index(typeid(T))) throw std::runtime_error("ndarray_group: wrong element type for " + name);
    return *static_cast<const ndarray<T>*>(e.ptr.get());
  }

 private:
  struct Entry {
    std::type_index type = std::type_index(typeid(void));
    std::shared_ptr<void> ptr;
  };
  std::map<std::string, Entry> arrays_;
};

// A sequence of timesteps read one at a time (total_timesteps(), read(t)).  The
// timesteps come from a callback; the YAML stream definitions of hguo/ndarray
// (files, synthetic generators) are not part of this header.
template <class = void>
class stream {
 public:
  using source = std::function<std::shared_ptr<ndarray_group>(int)>;
  stream() = default;
  stream(int n, source f) : n_(n), f_(std::move(f)) {}
  void set_source(int n, source f) {
    n_ = n;
    f_ = std::move(f);
  }
  int total_timesteps() const { return n_; }
  std::shared_ptr<ndarray_group> read(int t) const {
    if (!f_ || t < 0 || t >= n_) throw std::out_of_range("stream: no timestep " + std::to_string(t));
    return f_(t);
  }
  void parse_yaml(const std::string&) {
    throw std::runtime_error("stream: YAML stream definitions need hguo/ndarray; use set_source()");
  }

 private:
  int n_ = 0;
  source f_;
};

}  // namespace ftk
