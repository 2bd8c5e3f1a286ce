// FeatureComplex -> VTK XML files (.vtu UnstructuredGrid, .vtp PolyData) without
// the VTK library, with the arrays and cells of the reference writer
// (include/ftk2/utils/vtk.hpp:26-160 of hguo/ftk2):
//   points   float32, sum_i (double)barycentric_coords[0][i] * vertex coordinates
//            (first three coordinates; vtkPoints stores float)
//   Scalar   float32  FeatureElement::scalar
//   TrackID  int32    (int)FeatureElement::track_id
//   Type     int32    (int)FeatureElement::type
//   Time     float32  last coordinate when the mesh has more total than spatial
//            dimensions, else 0
//   .vtu     target_dim == 0: one VTK_VERTEX per node; otherwise the connectivity
//            lists of dimension 0..3 (all, or only target_dim) as VTK_VERTEX /
//            LINE / TRIANGLE / TETRA
//   .vtp     verts / lines / polys from connectivity dimensions 0 / 1 / 2
// Data are written inline as ASCII (use_ascii, or FTK_DEBUG set) or appended and
// base64-encoded without compression (header_type UInt64, little endian), the two
// modes the reference selects.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../core/complex.hpp"
#include "../core/mesh.hpp"

namespace ftk2 {
namespace vtk_detail {

struct PointArrays {
  std::vector<float> xyz, scalar, time;
  std::vector<int32_t> track, type;
};

inline PointArrays point_arrays(const FeatureComplex& complex, const Mesh& mesh) {
  PointArrays a;
  const size_t n = complex.vertices.size();
  const int td = mesh.get_total_dimension();
  const bool has_time = td > mesh.get_spatial_dimension();
  a.xyz.resize(3 * n);
  a.scalar.resize(n);
  a.time.resize(n);
  a.track.resize(n);
  a.type.resize(n);
  std::vector<double> pos(td);
  for (size_t k = 0; k < n; ++k) {
    const FeatureElement& el = complex.vertices[k];
    std::fill(pos.begin(), pos.end(), 0.0);
    for (int i = 0; i <= el.simplex.dimension; ++i) {
      const std::vector<double> c = mesh.get_vertex_coordinates(el.simplex.vertices[i]);
      for (int j = 0; j < td; ++j) pos[j] += (double)el.barycentric_coords[0][i] * c[j];
    }
    for (int j = 0; j < 3; ++j) a.xyz[3 * k + j] = j < td ? (float)pos[j] : 0.f;
    a.scalar[k] = el.scalar;
    a.track[k] = (int32_t)el.track_id;
    a.type[k] = (int32_t)el.type;
    a.time[k] = has_time ? (float)pos[td - 1] : 0.f;
  }
  return a;
}

inline void base64(const uint8_t* p, size_t n, std::string& out) {
  static const char* T = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
    out += T[v >> 18];
    out += T[(v >> 12) & 63];
    out += T[(v >> 6) & 63];
    out += T[v & 63];
  }
  if (i < n) {
    uint32_t v = (uint32_t)p[i] << 16;
    if (i + 1 < n) v |= (uint32_t)p[i + 1] << 8;
    out += T[v >> 18];
    out += T[(v >> 12) & 63];
    out += i + 1 < n ? T[(v >> 6) & 63] : '=';
    out += '=';
  }
}

// One data array: inline ASCII, or a reference into the appended block.
class Writer {
 public:
  explicit Writer(bool ascii) : ascii_(ascii) {}

  template <class V>
  void array(std::ostringstream& xml, const char* type, const char* name, int ncomp, const std::vector<V>& v) {
    xml << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\"";
    if (ncomp > 1) xml << " NumberOfComponents=\"" << ncomp << "\"";
    if (ascii_) {
      xml << " format=\"ascii\">\n          ";
      std::ostringstream vals;
      vals.precision(9);
      for (size_t i = 0; i < v.size(); ++i) {
        if (sizeof(V) == 1)
          vals << (int)v[i];
        else
          vals << v[i];
        vals << ((i + 1) % 12 == 0 ? "\n          " : " ");
      }
      xml << vals.str() << "\n        </DataArray>\n";
    } else {
      xml << " format=\"appended\" offset=\"" << appended_.size() << "\"/>\n";
      const uint64_t bytes = v.size() * sizeof(V);
      base64(reinterpret_cast<const uint8_t*>(&bytes), sizeof(bytes), appended_);
      base64(reinterpret_cast<const uint8_t*>(v.data()), bytes, appended_);
    }
  }

  void finish(std::ostringstream& xml) const {
    if (!ascii_) xml << "  <AppendedData encoding=\"base64\">\n   _" << appended_ << "\n  </AppendedData>\n";
    xml << "</VTKFile>\n";
  }

 private:
  bool ascii_;
  std::string appended_;
};

inline void header(std::ostringstream& xml, const char* type) {
  xml << "<?xml version=\"1.0\"?>\n<VTKFile type=\"" << type
      << "\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
}

inline void point_data(std::ostringstream& xml, Writer& w, const PointArrays& a) {
  xml << "      <PointData>\n";
  w.array(xml, "Float32", "Scalar", 1, a.scalar);
  w.array(xml, "Int32", "TrackID", 1, a.track);
  w.array(xml, "Int32", "Type", 1, a.type);
  w.array(xml, "Float32", "Time", 1, a.time);
  xml << "      </PointData>\n      <Points>\n";
  w.array(xml, "Float32", "Points", 3, a.xyz);
  xml << "      </Points>\n";
}

inline void save(const std::string& filename, const std::string& text) {
  std::ofstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + filename);
  f << text;
}

inline bool ascii_mode(bool use_ascii) { return use_ascii || std::getenv("FTK_DEBUG") != nullptr; }

}  // namespace vtk_detail

// write_complex_to_vtp (utils/vtk.hpp:78-113)
inline void write_complex_to_vtp(const FeatureComplex& complex, const Mesh& mesh, const std::string& filename,
                                 int target_dim = -1, bool use_ascii = false) {
  using namespace vtk_detail;
  const PointArrays a = point_arrays(complex, mesh);
  std::vector<int64_t> conn[3], offs[3];
  for (const auto& c : complex.connectivity) {
    if (target_dim != -1 && c.dimension != target_dim) continue;
    if (c.dimension < 0 || c.dimension > 2) continue;
    const int np = c.dimension + 1;
    for (size_t i = 0; i + np <= c.indices.size(); i += np) {
      for (int j = 0; j < np; ++j) conn[c.dimension].push_back(c.indices[i + j]);
      offs[c.dimension].push_back((int64_t)conn[c.dimension].size());
    }
  }
  std::ostringstream xml;
  header(xml, "PolyData");
  Writer w(ascii_mode(use_ascii));
  xml << "  <PolyData>\n    <Piece NumberOfPoints=\"" << complex.vertices.size() << "\" NumberOfVerts=\""
      << offs[0].size() << "\" NumberOfLines=\"" << offs[1].size() << "\" NumberOfStrips=\"0\" NumberOfPolys=\""
      << offs[2].size() << "\">\n";
  point_data(xml, w, a);
  const char* sect[3] = {"Verts", "Lines", "Polys"};
  for (int d = 0; d < 3; ++d) {
    xml << "      <" << sect[d] << ">\n";
    w.array(xml, "Int64", "connectivity", 1, conn[d]);
    w.array(xml, "Int64", "offsets", 1, offs[d]);
    xml << "      </" << sect[d] << ">\n";
  }
  xml << "    </Piece>\n  </PolyData>\n";
  w.finish(xml);
  save(filename, xml.str());
}

// write_complex_to_vtu (utils/vtk.hpp:118-160)
inline void write_complex_to_vtu(const FeatureComplex& complex, const Mesh& mesh, const std::string& filename,
                                 int target_dim = -1, bool use_ascii = false) {
  using namespace vtk_detail;
  const PointArrays a = point_arrays(complex, mesh);
  static const uint8_t vtk_type[4] = {1, 3, 5, 10};  // VERTEX, LINE, TRIANGLE, TETRA
  std::vector<int64_t> conn, offs;
  std::vector<uint8_t> types;
  if (target_dim == 0) {
    for (size_t i = 0; i < complex.vertices.size(); ++i) {
      conn.push_back((int64_t)i);
      offs.push_back((int64_t)conn.size());
      types.push_back(1);
    }
  } else {
    for (const auto& c : complex.connectivity) {
      if (target_dim != -1 && c.dimension != target_dim) continue;
      if (c.dimension < 0 || c.dimension > 3) continue;
      const int np = c.dimension + 1;
      for (size_t i = 0; i + np <= c.indices.size(); i += np) {
        for (int j = 0; j < np; ++j) conn.push_back(c.indices[i + j]);
        offs.push_back((int64_t)conn.size());
        types.push_back(vtk_type[c.dimension]);
      }
    }
  }
  std::ostringstream xml;
  header(xml, "UnstructuredGrid");
  Writer w(ascii_mode(use_ascii));
  xml << "  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" << complex.vertices.size() << "\" NumberOfCells=\""
      << types.size() << "\">\n";
  point_data(xml, w, a);
  xml << "      <Cells>\n";
  w.array(xml, "Int64", "connectivity", 1, conn);
  w.array(xml, "Int64", "offsets", 1, offs);
  w.array(xml, "UInt8", "types", 1, types);
  xml << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n";
  w.finish(xml);
  save(filename, xml.str());
}

}  // namespace ftk2
