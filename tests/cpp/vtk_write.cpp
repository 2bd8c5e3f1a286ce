// Writes a FeatureComplex read from a text dump through include/ftk2/utils/vtk.hpp.
// Input (whitespace separated):
//   d  global_dims[d]  m  n
//   n lines: vertices[m+1]  bary[m+1]  scalar  track_id
//   number of connectivity lists, then per list: dimension  cells  indices...
// Usage: vtk_write dump out_prefix   -> out_prefix{_ascii,_bin}.{vtu,vtp},
//                                      out_prefix_dim0.vtu, out_prefix_dim1.vtp
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "ftk2/utils/vtk.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::ifstream in(argv[1]);
  int d = 0, m = 0;
  size_t n = 0;
  in >> d;
  std::vector<uint64_t> dims(d);
  for (auto& x : dims) in >> x;
  in >> m >> n;
  ftk2::FeatureComplex cx;
  cx.vertices.resize(n);
  for (size_t k = 0; k < n; ++k) {
    ftk2::FeatureElement& el = cx.vertices[k];
    el.simplex.dimension = m;
    for (int i = 0; i <= m; ++i) in >> el.simplex.vertices[i];
    for (int i = 0; i <= m; ++i) in >> el.barycentric_coords[0][i];
    in >> el.scalar >> el.track_id;
  }
  int nl = 0;
  in >> nl;
  for (int l = 0; l < nl; ++l) {
    ftk2::FeatureComplex::SimplexIndices s;
    size_t cells = 0;
    in >> s.dimension >> cells;
    s.indices.resize(cells * (s.dimension + 1));
    for (auto& x : s.indices) in >> x;
    cx.connectivity.push_back(s);
  }
  if (!in) {
    std::fprintf(stderr, "bad dump\n");
    return 1;
  }
  ftk2::RegularSimplicialMesh mesh(dims);
  const std::string p = argv[2];
  ftk2::write_complex_to_vtu(cx, mesh, p + "_ascii.vtu", -1, true);
  ftk2::write_complex_to_vtu(cx, mesh, p + "_bin.vtu", -1, false);
  ftk2::write_complex_to_vtp(cx, mesh, p + "_ascii.vtp", -1, true);
  ftk2::write_complex_to_vtp(cx, mesh, p + "_bin.vtp", -1, false);
  ftk2::write_complex_to_vtu(cx, mesh, p + "_dim0.vtu", 0, false);
  ftk2::write_complex_to_vtp(cx, mesh, p + "_dim1.vtp", 1, false);
  std::printf("wrote %zu nodes\n", n);
  return 0;
}
