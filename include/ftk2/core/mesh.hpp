// Simplex and RegularSimplicialMesh of FTK2's public API
// (hguo/ftk2 include/ftk2/core/mesh.hpp:31-65, :80-103, :256-537), the part that
// callers of SimplicialEngine construct and read.  Enumeration, coface queries and
// predicate evaluation happen on the GPU behind include/ftk2h.h.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace ftk2 {

struct Simplex {  // 48 bytes, as in the reference
  int dimension = 0;
  uint64_t vertices[5] = {0, 0, 0, 0, 0};
  bool operator<(const Simplex& o) const {
    if (dimension != o.dimension) return dimension < o.dimension;
    for (int i = 0; i <= dimension; ++i)
      if (vertices[i] != o.vertices[i]) return vertices[i] < o.vertices[i];
    return false;
  }
  bool operator==(const Simplex& o) const {
    if (dimension != o.dimension) return false;
    for (int i = 0; i <= dimension; ++i)
      if (vertices[i] != o.vertices[i]) return false;
    return true;
  }
};

class Mesh {
 public:
  virtual ~Mesh() = default;
  virtual int get_spatial_dimension() const = 0;
  virtual int get_total_dimension() const = 0;
  virtual uint64_t get_num_vertices() const = 0;
  virtual std::vector<double> get_vertex_coordinates(uint64_t vertex_id) const = 0;
};

// Regular grid whose last axis is time; vertex id = x + gx*(y + gy*(z + gz*t)) over
// global extents of (local + offset) coordinates.
class RegularSimplicialMesh : public Mesh {
 public:
  RegularSimplicialMesh(const std::vector<uint64_t>& dims, const std::vector<uint64_t>& offset = {},
                        const std::vector<uint64_t>& global_dims = {})
      : dims_(dims), offset_(offset), global_dims_(global_dims) {
    if (offset_.empty()) offset_.assign(dims_.size(), 0);
    if (global_dims_.empty()) global_dims_ = dims_;
  }
  int get_spatial_dimension() const override { return (int)dims_.size(); }
  int get_total_dimension() const override { return (int)dims_.size(); }
  uint64_t get_num_vertices() const override {
    uint64_t n = 1;
    for (auto d : dims_) n *= d;
    return n;
  }
  std::vector<double> get_vertex_coordinates(uint64_t id) const override {
    std::vector<double> c(global_dims_.size());
    for (size_t i = 0; i < global_dims_.size(); ++i) {
      c[i] = (double)(id % global_dims_[i]);
      id /= global_dims_[i];
    }
    return c;
  }
  const std::vector<uint64_t>& get_local_dims() const { return dims_; }
  const std::vector<uint64_t>& get_offset() const { return offset_; }
  const std::vector<uint64_t>& get_global_dims() const { return global_dims_; }

 private:
  std::vector<uint64_t> dims_, offset_, global_dims_;
};

}  // namespace ftk2
