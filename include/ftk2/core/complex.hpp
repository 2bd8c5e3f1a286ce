// FeatureComplex (hguo/ftk2 include/ftk2/core/complex.hpp:14-58).  Complexes
// produced by SimplicialEngine::get_complex() are already in sort() order.
#pragma once
#include <algorithm>
#include <vector>

#include "feature.hpp"

namespace ftk2 {

struct FeatureComplex {
  std::vector<FeatureElement> vertices;
  struct SimplexIndices {
    int dimension;
    std::vector<uint32_t> indices;
  };
  std::vector<SimplexIndices> connectivity;

  // vertices by Simplex; cells with sorted indices, lexicographic, unique
  void sort() {
    std::sort(vertices.begin(), vertices.end(),
              [](const FeatureElement& a, const FeatureElement& b) { return a.simplex < b.simplex; });
    std::sort(connectivity.begin(), connectivity.end(),
              [](const SimplexIndices& a, const SimplexIndices& b) { return a.dimension < b.dimension; });
    for (auto& conn : connectivity) {
      const int n = conn.dimension + 1;
      std::vector<std::vector<uint32_t>> cells;
      for (size_t i = 0; i + n <= conn.indices.size(); i += n) {
        std::vector<uint32_t> c(conn.indices.begin() + i, conn.indices.begin() + i + n);
        std::sort(c.begin(), c.end());
        cells.push_back(c);
      }
      std::sort(cells.begin(), cells.end());
      cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
      conn.indices.clear();
      for (const auto& c : cells) conn.indices.insert(conn.indices.end(), c.begin(), c.end());
    }
  }
};

}  // namespace ftk2
