// FeatureElement: the 328-byte record of FTK2 (hguo/ftk2
// include/ftk2/core/feature.hpp:22-39), same field order and types.
#pragma once
#include <cstring>

#include "mesh.hpp"

namespace ftk2 {

enum class FeatureGeometryType { Point, Segment, Polygon };

struct FeatureElement {
  Simplex simplex;
  FeatureGeometryType geometry_type;
  float barycentric_coords[3][16];
  uint64_t track_id;
  uint32_t type;
  float scalar;
  float attributes[16];
  FeatureElement() {
    std::memset(static_cast<void*>(this), 0, sizeof(FeatureElement));
    geometry_type = FeatureGeometryType::Point;
  }
};

}  // namespace ftk2
