// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own headers (hguo/ftk2
// include/ftk2/core/sos.hpp, mesh.hpp with coface_lut.hpp and parallel.hpp), which
// compile as they are.  oracle/Makefile (target `ref`) builds this file against the
// reference tree into oracle/_ref/libftk2ref.so; tests/test_oracle_ref.py checks the
// oracle's restatement against it: the sign tests (sos.hpp:12-254, all det3 levels),
// the simplex LUTs and their enumeration (mesh.hpp:111-302), encode_simplex_id
// (mesh.hpp:736-749) and the cofaces pass of get_complex (engine.hpp:896-963, with
// Mesh::cofaces / faces, mesh.hpp:90-102,304-372).  The engine headers
// (engine.hpp, predicate.hpp, zero_crossing.hpp through feature.hpp) need hguo/ndarray,
// which is not available, and are not built.
#include <ftk2/core/sos.hpp>

#include <cstdint>
#include <vector>

using ftk2::RegularSimplicialMesh;
using ftk2::Simplex;

namespace {
template <class T>
void load(T (&dst)[5][4], const double* src, int rows, int cols) {
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) dst[i][j] = (T)src[i * cols + j];
}
std::vector<uint64_t> vec(const uint64_t* p, int n) { return std::vector<uint64_t>(p, p + n); }
}  // namespace

extern "C" {

int ref_sign(int dtype, double v, uint64_t idx, double Q) {
  return dtype == 0 ? ftk2::sos::sign<float>((float)v, idx, Q) : ftk2::sos::sign<double>(v, idx, Q);
}

int ref_det2(int dtype, const double* a, const double* b, uint64_t i0, uint64_t i1, double Q) {
  if (dtype == 0) {
    float x[2] = {(float)a[0], (float)a[1]}, y[2] = {(float)b[0], (float)b[1]};
    return ftk2::sos::det2(x, y, i0, i1, Q);
  }
  return ftk2::sos::det2(a, b, i0, i1, Q);
}

int ref_det3(int dtype, const double* a, const double* b, const double* c, double Q) {
  if (dtype == 0) {
    float x[3], y[3], z[3];
    for (int j = 0; j < 3; ++j) {
      x[j] = (float)a[j];
      y[j] = (float)b[j];
      z[j] = (float)c[j];
    }
    return ftk2::sos::det3(x, y, z, 0, 1, 2, Q);
  }
  return ftk2::sos::det3(a, b, c, 0, 1, 2, Q);
}

// origin_inside<K> on n simplices at once: vals[n][K+1][K], idx[n][K+1] -> out[n]
void ref_origin_inside_many(int K, int dtype, int64_t n, const double* vals, const uint64_t* idx, double Q,
                            int* out) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    for (int64_t s = 0; s < n; ++s) {
      T v[5][4];
      load(v, vals + s * (K + 1) * K, K + 1, K);
      const uint64_t* ix = idx + s * (K + 1);
      int r = -1;
      if (K == 1) {
        T w[2][1] = {{v[0][0]}, {v[1][0]}};
        r = ftk2::sos::origin_inside<1, T>::check(w, ix, Q);
      } else if (K == 2) {
        T w[3][2];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 2; ++j) w[i][j] = v[i][j];
        r = ftk2::sos::origin_inside<2, T>::check(w, ix, Q);
      } else if (K == 3) {
        T w[4][3];
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 3; ++j) w[i][j] = v[i][j];
        r = ftk2::sos::origin_inside<3, T>::check(w, ix, Q);
      }
      out[s] = r;
    }
  };
  if (dtype == 0)
    go(float{});
  else
    go(double{});
}

int ref_num_simplex_types(int d, int m) { return ftk2::cpu_get_num_simplex_types(d, m); }

int ref_lut(int d, int m, int type, int vi, int j) { return ftk2::cpu_lut_dispatch(d, m, type, vi, j); }

// RegularSimplicialMesh::iterate_simplices(k) on a mesh of local dims `dims` (offset
// and global dims optional): writes up to cap simplices of k+1 sorted ids, returns the count.
uint64_t ref_iterate_simplices(int d, const uint64_t* dims, const uint64_t* off, const uint64_t* gdims, int k,
                               uint64_t* out, uint64_t cap) {
  RegularSimplicialMesh mesh(vec(dims, d), off ? vec(off, d) : std::vector<uint64_t>{},
                             gdims ? vec(gdims, d) : std::vector<uint64_t>{});
  uint64_t n = 0;
  mesh.iterate_simplices(k, [&](const Simplex& s) {
    if (n < cap)
      for (int i = 0; i <= k; ++i) out[n * (k + 1) + i] = s.vertices[i];
    ++n;
  });
  return n;
}

uint64_t ref_encode_simplex_id(int d, const uint64_t* gdims, int dim, const uint64_t* verts) {
  RegularSimplicialMesh mesh(vec(gdims, d));
  Simplex s;
  s.dimension = dim;
  for (int i = 0; i <= dim; ++i) s.vertices[i] = verts[i];
  return ftk2::encode_simplex_id(s, mesh);
}

// The neighbours the cofaces pass of get_complex() (engine.hpp:896-963) unites a
// tetrahedron with on a 4-D mesh: every tet face of every pentatope coface of it
// (Mesh::cofaces, then Mesh::faces).  Writes up to cap sorted 4-id tets, returns the
// count (with repetitions, as the pass visits them before its std::set).
int ref_coface_neighbors(const uint64_t* dims, const uint64_t* off, const uint64_t* gdims, const uint64_t* tet,
                         uint64_t* out, int cap) {
  RegularSimplicialMesh mesh(vec(dims, 4), off ? vec(off, 4) : std::vector<uint64_t>{},
                             gdims ? vec(gdims, 4) : std::vector<uint64_t>{});
  Simplex s;
  s.dimension = 3;
  for (int i = 0; i < 4; ++i) s.vertices[i] = tet[i];
  int n = 0;
  mesh.cofaces(s, [&](const Simplex& pent) {
    if (pent.dimension != 4) return;
    mesh.faces(pent, [&](const Simplex& f) {
      if (f.dimension != 3) return;
      Simplex t = f;
      t.sort_vertices();
      if (n < cap)
        for (int i = 0; i < 4; ++i) out[n * 4 + i] = t.vertices[i];
      ++n;
    });
  });
  return n;
}

}  // extern "C"
