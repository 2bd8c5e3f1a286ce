// Simplex and RegularSimplicialMesh of FTK2's public API
// (hguo/ftk2 include/ftk2/core/mesh.hpp:31-65, :80-103, :256-537) and
// ExtrudedSimplicialMesh over a regular base (:542-726), the part that
// callers of SimplicialEngine construct and read.  Enumeration, coface queries and
// predicate evaluation happen on the GPU behind include/ftk2h.h.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
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
  // grid_index_to_id (mesh.hpp:439-446): sum_j g[j] * prod_{i<j} global_dims[i]
  uint64_t grid_index_to_id(const uint64_t g[4]) const {
    uint64_t id = 0, mul = 1;
    for (size_t i = 0; i < global_dims_.size(); ++i) {
      id += g[i] * mul;
      mul *= global_dims_[i];
    }
    return id;
  }

 private:
  std::vector<uint64_t> dims_, offset_, global_dims_;
};

// A base mesh stacked in n_layers + 1 time layers (mesh.hpp:542-560,711-726):
// vertex id = spatial id + t * n_spatial_verts_, where for a regular base
// n_spatial_verts_ = grid_index_to_id(base global dims), which exceeds the number
// of spatial vertices.  Its simplices are the Kuhn simplices of the base grid with
// one more axis, so SimplicialEngine runs it on the regular path with these ids.
class ExtrudedSimplicialMesh : public Mesh {
 public:
  ExtrudedSimplicialMesh(std::shared_ptr<Mesh> base_mesh, uint64_t n_layers = 0)
      : base_mesh_(std::move(base_mesh)), n_layers_(n_layers) {
    auto reg = std::dynamic_pointer_cast<RegularSimplicialMesh>(base_mesh_);
    if (!reg) throw std::runtime_error("ExtrudedSimplicialMesh: only a RegularSimplicialMesh base is supported");
    uint64_t g[4] = {0, 0, 0, 0};
    const auto& gd = reg->get_global_dims();
    for (size_t i = 0; i < 4 && i < gd.size(); ++i) g[i] = gd[i];
    n_spatial_verts_ = reg->grid_index_to_id(g);
  }
  int get_spatial_dimension() const override { return base_mesh_->get_spatial_dimension(); }
  int get_total_dimension() const override { return base_mesh_->get_total_dimension() + 1; }
  uint64_t get_num_vertices() const override { return n_spatial_verts_ * (n_layers_ + 1); }
  std::vector<double> get_vertex_coordinates(uint64_t vertex_id) const override {
    std::vector<double> c = base_mesh_->get_vertex_coordinates(vertex_id % n_spatial_verts_);
    c.push_back((double)(vertex_id / n_spatial_verts_));
    return c;
  }
  std::shared_ptr<Mesh> get_base_mesh() const { return base_mesh_; }
  uint64_t get_n_layers() const { return n_layers_; }
  uint64_t get_n_spatial_verts() const { return n_spatial_verts_; }

 private:
  std::shared_ptr<Mesh> base_mesh_;
  uint64_t n_layers_;
  uint64_t n_spatial_verts_ = 0;
};

}  // namespace ftk2
