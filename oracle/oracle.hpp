// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement of FTK2's CPU simplicial engine
// (include/ftk2/core/{mesh,sos,zero_crossing,predicate,engine,complex}.hpp).
// It exists to check the HIP implementation: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only as the checker or the
// timed CPU baseline.  The product path (ftk2_amd/, include/) never calls it.
//
// Every routine cites the reference function it restates.  Where the reference
// result does not depend on an internal detail (node ids handed out in discovery
// order, a global mutex, std::map containers) the restatement uses plain sorted
// vectors instead; the observable output (FeatureComplex after sort()) is the same
// because get_complex() re-numbers nodes by Simplex order and sorts every cell list.
//
// Parity pins (see tests/test_oracle_*.py):
//   sign/det2/det3/origin_inside  : test_sos.cpp known answers, exhaustive small-int
//                                   boxes against exact rational geometry, level-0 =
//                                   triple product, d_i = -V_i.
//   solver                        : test_zero_crossing.cpp / test_3d_cp.cpp answers,
//                                   numpy solve within 1e-12, exact 0.25 case.
//   engine end to end             : test_examples.cpp golden counts 6078 / 66418 / 788
//                                   (the last on a restated woven field),
//                                   test_engine.cpp structure, test_mesh.cpp counts,
//                                   brute-force "every m-face of every d-cell" node set,
//                                   union-find against scipy connected components.
//                                   The CP3D float case asserted as 52 there
//                                   (test_examples.cpp:153-186) gives 29 here, as a
//                                   literal transcription of the reference does.
// Parity unpinned: d = 3 contour connectivity where a 2-2 split picks its quad
// diagonal (the reference uses node discovery order, engine.hpp:1335-1337), and the
// levelset-2D golden 1206, whose input generator is not available.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

using i128 = __int128;

// Simplex: mesh.hpp:31-65 (dimension, then vertex ids lexicographically).
struct Simplex {
  int dimension = 0;
  uint64_t v[5] = {0, 0, 0, 0, 0};
  bool operator<(const Simplex& o) const {
    if (dimension != o.dimension) return dimension < o.dimension;
    for (int i = 0; i <= dimension; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return false;
  }
  bool operator==(const Simplex& o) const {
    if (dimension != o.dimension) return false;
    for (int i = 0; i <= dimension; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
  void sort_vertices() { std::sort(v, v + dimension + 1); }  // mesh.hpp:54-64
};

// ---------------------------------------------------------------------------
// Freudenthal LUTs (mesh.hpp:111-236).  Every table row of the reference is a
// chain 0 = s0 < s1 < ... < sm of vertex offsets in {0,1}^d (strict inclusion of
// the set of raised axes); the reference lists each chain once.  We enumerate the
// chains directly; tests check the counts 3,2,7,12,6,15,50,60,24.
struct Chains {
  int d = 0, m = 0, n = 0;
  uint8_t s[64][5];  // s[type][i]: bit j set <=> axis j raised at vertex i
};

inline void chains_rec(Chains& c, int depth, uint8_t cur, uint8_t* stack) {
  if (depth == c.m) {
    for (int i = 0; i <= c.m; ++i) c.s[c.n][i] = stack[i];
    c.n++;
    return;
  }
  const uint8_t full = (uint8_t)((1u << c.d) - 1);
  for (uint8_t nx = 1; nx <= full; ++nx) {
    if ((nx & cur) != cur || nx == cur) continue;  // strict superset of cur
    stack[depth + 1] = nx;
    chains_rec(c, depth + 1, nx, stack);
  }
}

inline Chains make_chains(int d, int m) {
  Chains c;
  c.d = d;
  c.m = m;
  c.n = 0;
  if (d < 1 || d > 4 || m < 1 || m > d) return c;
  uint8_t stack[5] = {0, 0, 0, 0, 0};
  chains_rec(c, 0, 0, stack);
  return c;
}

// RegularSimplicialMesh (mesh.hpp:256-537): local dims, offset, global dims;
// vertex id = x + gx*(y + gy*(z + gz*t)) over GLOBAL dims (mesh.hpp:439-451).
struct Mesh {
  int d = 0;
  uint64_t dims[4] = {1, 1, 1, 1}, off[4] = {0, 0, 0, 0}, gdims[4] = {1, 1, 1, 1};
  // Nonzero: ExtrudedSimplicialMesh over a regular base (mesh.hpp:542-560,711-717):
  // the last axis is the layer t and id = spatial id + t * tstride, where tstride
  // is the base's n_spatial_verts_ = grid_index_to_id(global dims of the base).
  uint64_t tstride = 0;
  uint64_t id(const uint64_t g[4]) const {
    uint64_t r = 0, mul = 1;
    for (int j = 0; j < d; ++j) {
      r += g[j] * (tstride && j == d - 1 ? tstride : mul);
      mul *= gdims[j];
    }
    return r;
  }
  void coords(uint64_t id_, uint64_t g[4]) const {  // mesh.hpp:420-424
    for (int j = 0; j < 4; ++j) g[j] = 0;
    int sd = d;
    if (tstride) {  // t = id / n_spatial_verts_, spatial = id % n_spatial_verts_
      g[d - 1] = id_ / tstride;
      id_ %= tstride;
      sd = d - 1;
    }
    for (int j = 0; j < sd; ++j) {
      g[j] = id_ % gdims[j];
      id_ /= gdims[j];
    }
  }
  uint64_t n_vertices() const {
    uint64_t n = 1;
    for (int j = 0; j < d; ++j) n *= dims[j];
    return n;
  }
  bool has_hypercubes() const {
    for (int j = 0; j < d; ++j)
      if (dims[j] < 2) return false;
    return true;
  }
  uint64_t n_hypercubes() const {
    if (!has_hypercubes()) return 0;
    uint64_t n = 1;
    for (int j = 0; j < d; ++j) n *= dims[j] - 1;
    return n;
  }
  // hypercube index -> local base coords (mesh.hpp:403-406, x fastest)
  void hc_base(uint64_t hc, uint64_t b[4]) const {
    for (int j = 0; j < 4; ++j) b[j] = 0;
    for (int j = 0; j < d; ++j) {
      b[j] = hc % (dims[j] - 1);
      hc /= (dims[j] - 1);
    }
  }
};

// encode_simplex_id (mesh.hpp:736-749): (v_min << 24) | 4-bit "coord > c_min"
// masks of the other vertices, in vertex order.
inline uint64_t encode_simplex_id(const Simplex& s, const Mesh& mesh) {
  uint64_t vmin = s.v[0];
  for (int i = 1; i <= s.dimension; ++i) vmin = std::min(vmin, s.v[i]);
  uint64_t cmin[4];
  mesh.coords(vmin, cmin);
  uint64_t comb = 0;
  for (int i = 0; i <= s.dimension; ++i) {
    if (s.v[i] == vmin) continue;
    uint64_t ci[4];
    mesh.coords(s.v[i], ci);
    uint64_t mask = 0;
    for (int k = 0; k < 4 && k < mesh.d; ++k)
      if (ci[k] > cmin[k]) mask |= (1ull << k);
    comb = (comb << 4) | (mask & 0xF);
  }
  return (vmin << 24) | (comb & 0xFFFFFF);
}

// ---------------------------------------------------------------------------
// Exact sign tests (sos.hpp).
namespace sos {

// Fixed-point quantization used inside every sos routine (sos.hpp:15,25,75):
// round half away from zero after two IEEE double operations (no FMA).
template <class T>
inline i128 q(T val, double Q) {
  double prod = static_cast<double>(val) * Q;  // separate rounding: -ffp-contract=off
  double r = prod + (val > 0 ? 0.5 : -0.5);
  return static_cast<i128>(r);
}

// sos.hpp:12-20
template <class T>
inline int sign(T v, uint64_t index, double Q) {
  i128 qv = q(v, Q);
  if (qv > 0) return 1;
  if (qv < 0) return -1;
  return (index % 2 == 0) ? 1 : -1;
}

// sos.hpp:22-35
template <class T>
inline int det2(const T v0[2], const T v1[2], uint64_t i0, uint64_t i1, double Q) {
  i128 a = q(v0[0], Q), b = q(v0[1], Q), c = q(v1[0], Q), d = q(v1[1], Q);
  i128 det = a * d - b * c;
  if (det > 0) return 1;
  if (det < 0) return -1;
  if (i0 < i1) return 1;
  if (i0 > i1) return -1;
  return 0;
}

// Exact determinant of a small integer matrix (n <= 4) by cofactor expansion.
inline i128 idet(const i128* M, int n) {
  if (n == 1) return M[0];
  if (n == 2) return M[0] * M[3] - M[1] * M[2];
  i128 r = 0;
  i128 sub[16];
  for (int c = 0; c < n; ++c) {
    int k = 0;
    for (int i = 1; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (j != c) sub[k++] = M[i * n + j];
    i128 cof = idet(sub, n - 1);
    r += ((c & 1) ? -M[c] : M[c]) * cof;
  }
  return r;
}

// The 15 perturbation levels of sos.hpp:72-200, restated as a table.  Rows index
// X[0..3] = (v0, v1, v2, origin); columns are vector components; every minor
// carries an extra all-ones column on the right; `neg` flips the sign.  Level 0
// is the 4x4 determinant over all components.  Level 14 returns +1.  The index
// arguments of the reference det3 are unused, and so are they here.
struct Level {
  int nrows;
  int rows[4];
  int ncols;
  int cols[3];
  bool neg;
};
inline const Level* det3_levels() {
  static const Level L[14] = {
      {4, {0, 1, 2, 3}, 3, {0, 1, 2}, false},  // t=0
      {3, {1, 2, 3}, 2, {0, 1}, false},        // t=1
      {3, {1, 2, 3}, 2, {0, 2}, true},         // t=2
      {3, {1, 2, 3}, 2, {1, 2}, false},        // t=3
      {3, {0, 2, 3}, 2, {0, 1}, true},         // t=4
      {2, {2, 3}, 1, {0}, false},              // t=5
      {2, {2, 3}, 1, {1}, true},               // t=6
      {3, {0, 2, 3}, 2, {0, 2}, false},        // t=7
      {2, {2, 3}, 1, {2}, false},              // t=8
      {3, {0, 2, 3}, 2, {1, 2}, true},         // t=9
      {3, {0, 1, 3}, 2, {0, 1}, false},        // t=10
      {2, {1, 3}, 1, {0}, true},               // t=11
      {2, {1, 3}, 1, {1}, false},              // t=12
      {2, {0, 3}, 1, {0}, false},              // t=13
  };
  return L;
}

template <class T>
inline int det3(const T v0[3], const T v1[3], const T v2[3], double Q) {
  i128 X[4][3];
  for (int j = 0; j < 3; ++j) {
    X[0][j] = q(v0[j], Q);
    X[1][j] = q(v1[j], Q);
    X[2][j] = q(v2[j], Q);
    X[3][j] = q(T(0), Q);  // the origin row
  }
  const Level* L = det3_levels();
  for (int t = 0; t < 14; ++t) {
    const Level& lv = L[t];
    int n = lv.nrows;
    i128 M[16];
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < lv.ncols; ++j) M[i * n + j] = X[lv.rows[i]][lv.cols[j]];
      M[i * n + lv.ncols] = 1;
    }
    i128 det = idet(M, n);
    if (lv.neg) det = -det;
    if (det > 0) return 1;
    if (det < 0) return -1;
  }
  return 1;  // t == 14
}

// origin_inside<K> (sos.hpp:202-254); vertices are in ascending id order.
template <int K, class T>
struct origin_inside;
template <class T>
struct origin_inside<1, T> {
  static bool check(const T v[2][1], const uint64_t idx[2], double Q) {
    return sign(v[0][0], idx[0], Q) != sign(v[1][0], idx[1], Q);
  }
};
template <class T>
struct origin_inside<2, T> {
  static bool check(const T v[3][2], const uint64_t idx[3], double Q) {
    int s0 = det2(v[0], v[1], idx[0], idx[1], Q);
    int s1 = det2(v[1], v[2], idx[1], idx[2], Q);
    int s2 = det2(v[2], v[0], idx[2], idx[0], Q);
    return s0 == s1 && s1 == s2;
  }
};
template <class T>
struct origin_inside<3, T> {
  static bool check(const T v[4][3], const uint64_t idx[4], double Q) {
    (void)idx;
    int d0 = det3(v[1], v[3], v[2], Q);
    int d1 = det3(v[0], v[2], v[3], Q);
    int d2 = det3(v[0], v[3], v[1], Q);
    int d3 = det3(v[0], v[1], v[2], Q);
    return d0 == d1 && d1 == d2 && d2 == d3;
  }
};

}  // namespace sos

// ---------------------------------------------------------------------------
// ZeroCrossingSolver (zero_crossing.hpp:20-105).  Gaussian elimination with
// first-maximum partial pivoting; every product and sum is rounded on its own
// (the reference's x86-64 build has no FMA).  Build with -ffp-contract=off.
template <int K, class T>
struct Solver {
  static bool solve(const T values[K + 1][K], T lambda[K + 1]) {
    T A[K][K], b[K];
    for (int j = 0; j < K; ++j) {
      for (int i = 0; i < K; ++i) A[j][i] = values[i][j] - values[K][j];
      b[j] = -values[K][j];
    }
    T x[K];
    if (!gepp(A, b, x)) return false;
    T sum = 0;
    for (int i = 0; i < K; ++i) {
      lambda[i] = x[i];
      sum += lambda[i];
    }
    lambda[K] = T(1.0) - sum;
    const T eps = T(1e-9);
    for (int i = 0; i <= K; ++i)
      if (lambda[i] < -eps || lambda[i] > T(1.0) + eps) return false;
    return true;
  }
  static bool gepp(T A[K][K], const T b[K], T x[K]) {
    T a[K][K];
    for (int i = 0; i < K; ++i) {
      for (int j = 0; j < K; ++j) a[i][j] = A[i][j];
      x[i] = b[i];
    }
    for (int i = 0; i < K; ++i) {
      int p = i;
      for (int j = i + 1; j < K; ++j)
        if (std::abs(a[j][i]) > std::abs(a[p][i])) p = j;
      for (int l = 0; l < K; ++l) std::swap(a[i][l], a[p][l]);
      std::swap(x[i], x[p]);
      if (std::abs(a[i][i]) < T(1e-12)) return false;
      for (int j = i + 1; j < K; ++j) {
        T f = a[j][i] / a[i][i];
        x[j] -= f * x[i];
        for (int l = i + 1; l < K; ++l) a[j][l] -= f * a[i][l];
      }
    }
    for (int i = K - 1; i >= 0; --i) {
      T s = 0;
      for (int j = i + 1; j < K; ++j) s += a[i][j] * x[j];
      x[i] = (x[i] - s) / a[i][i];
    }
    return true;
  }
};
template <class T>
struct Solver<1, T> {
  static bool solve(const T values[2][1], T lambda[2]) {
    T d = values[0][0] - values[1][0];
    if (std::abs(d) < T(1e-12)) return false;
    lambda[0] = -values[1][0] / d;
    lambda[1] = T(1.0) - lambda[0];
    const T eps = T(1e-9);
    return lambda[0] >= -eps && lambda[0] <= T(1.0) + eps;
  }
};

// ---------------------------------------------------------------------------
// Predicates (predicate.hpp) and engine (engine.hpp).
enum PredKind { PRED_CP = 0, PRED_CONTOUR = 1, PRED_FIBER = 2 };

// A field component as the engine reads it (engine.hpp:999-1009 get_value):
// element at LOCAL coordinates (global - offset), element strides per axis.
template <class T>
struct View {
  const T* p = nullptr;
  int64_t st[4] = {0, 0, 0, 0};
  T at(const uint64_t loc[4], int d) const {
    int64_t i = 0;
    for (int j = 0; j < d; ++j) i += (int64_t)loc[j] * st[j];
    return p[i];
  }
};

template <class T>
struct Params {
  int kind = PRED_CP;
  int m = 3;  // codimension
  double Q = 1000000.0;
  T thr[2] = {T(0), T(0)};
  int ncomp = 0;
  View<T> comp[3];
  // CriticalPointPredicate scalar interpolation (predicate.hpp:81-92): legacy mode
  // only, read at GLOBAL vertex coordinates.
  bool has_scalar = false;
  View<T> scalar;
};

struct Node {
  Simplex s;
  double lambda[5];  // solver output (T widened), before the float cast
  float bary[5];     // FeatureElement::barycentric_coords[0][i]
  float scalar;
  bool operator<(const Node& o) const { return s < o.s; }
};

// One predicate evaluation on sorted simplex s (engine.hpp:1036-1204 extract_simplex
// + predicate.hpp extract_it).
template <class T>
inline bool extract(const Params<T>& P, const Mesh& mesh, const Simplex& s, Node& out) {
  const int nv = s.dimension + 1;
  uint64_t loc[5][4];
  for (int i = 0; i < nv; ++i) {
    uint64_t g[4];
    mesh.coords(s.v[i], g);
    for (int j = 0; j < 4; ++j) loc[i][j] = g[j] - (j < mesh.d ? mesh.off[j] : 0);
  }
  out.s = s;
  for (int i = 0; i < 5; ++i) {
    out.lambda[i] = 0;
    out.bary[i] = 0.f;
  }
  out.scalar = 0.f;
  const double Q = P.Q;
  if (P.kind == PRED_CONTOUR) {  // predicate.hpp:143-156
    if (nv != 2) return false;
    T adj[2][1];
    for (int i = 0; i < 2; ++i) adj[i][0] = P.comp[0].at(loc[i], mesh.d) - P.thr[0];
    uint64_t idx[2] = {s.v[0], s.v[1]};
    if (!sos::origin_inside<1, T>::check(adj, idx, Q)) return false;
    T lam[2];
    if (!Solver<1, T>::solve(adj, lam)) lam[0] = lam[1] = (T)0.5;
    for (int i = 0; i < 2; ++i) {
      out.lambda[i] = (double)lam[i];
      out.bary[i] = (float)lam[i];
    }
    out.scalar = (float)P.thr[0];
    return true;
  }
  if (P.kind == PRED_FIBER) {  // predicate.hpp:348-362
    if (nv != 3) return false;
    T adj[3][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 2; ++j) adj[i][j] = P.comp[j].at(loc[i], mesh.d) - P.thr[j];
    uint64_t idx[3] = {s.v[0], s.v[1], s.v[2]};
    if (!sos::origin_inside<2, T>::check(adj, idx, Q)) return false;
    T lam[3];
    if (!Solver<2, T>::solve(adj, lam))
      for (int i = 0; i < 3; ++i) lam[i] = (T)1.0 / 3.0;
    for (int i = 0; i < 3; ++i) {
      out.lambda[i] = (double)lam[i];
      out.bary[i] = (float)lam[i];
    }
    return true;
  }
  // CriticalPointPredicate<M> (predicate.hpp:66-99)
  const int M = P.m;
  T lam[4];
  bool ok = false;
  if (M == 1) {
    T v[2][1];
    for (int i = 0; i < 2; ++i) v[i][0] = P.comp[0].at(loc[i], mesh.d);
    uint64_t idx[2] = {s.v[0], s.v[1]};
    ok = sos::origin_inside<1, T>::check(v, idx, Q) && Solver<1, T>::solve(v, lam);
  } else if (M == 2) {
    T v[3][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 2; ++j) v[i][j] = P.comp[j].at(loc[i], mesh.d);
    uint64_t idx[3] = {s.v[0], s.v[1], s.v[2]};
    ok = sos::origin_inside<2, T>::check(v, idx, Q) && Solver<2, T>::solve(v, lam);
  } else if (M == 3) {
    T v[4][3];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) v[i][j] = P.comp[j].at(loc[i], mesh.d);
    uint64_t idx[4] = {s.v[0], s.v[1], s.v[2], s.v[3]};
    ok = sos::origin_inside<3, T>::check(v, idx, Q) && Solver<3, T>::solve(v, lam);
  }
  if (!ok) return false;
  for (int i = 0; i <= M; ++i) {
    out.lambda[i] = (double)lam[i];
    out.bary[i] = (float)lam[i];
  }
  if (P.has_scalar) {  // predicate.hpp:81-92, global coordinates
    T sv = 0;
    for (int i = 0; i <= M; ++i) {
      uint64_t g[4];
      mesh.coords(s.v[i], g);
      sv += lam[i] * P.scalar.at(g, mesh.d);
    }
    out.scalar = (float)sv;
  }
  return true;
}

// Kuhn d-cell (mesh.hpp:403-414): permutation decoded by successive swaps.
inline Simplex d_cell(const Mesh& mesh, const uint64_t base_local[4], int p_idx) {
  const int d = mesh.d;
  int p[4] = {0, 1, 2, 3};
  int t = p_idx;
  for (int i = 0; i < d; ++i) {
    int j = t % (d - i);
    std::swap(p[i], p[i + j]);
    t /= (d - i);
  }
  uint64_t cur[4] = {0, 0, 0, 0};
  for (int j = 0; j < d; ++j) cur[j] = base_local[j] + mesh.off[j];
  Simplex s;
  s.dimension = d;
  s.v[0] = mesh.id(cur);
  for (int i = 0; i < d; ++i) {
    cur[p[i]]++;
    s.v[i + 1] = mesh.id(cur);
  }
  s.sort_vertices();
  return s;
}

// All results of one engine (active nodes + manifold simplices), keyed by Simplex.
struct Accum {
  std::vector<Node> nodes;                       // unique, sorted after finalize
  std::vector<std::vector<Simplex>> manifold[4];  // [dim] -> cells as node simplices
};

// feed() for one (slab) mesh: LUT node pass (engine.hpp:118-158, mesh.hpp:270-302)
// and the manifold pass over every d-cell (engine.hpp:161-206): star edges / star-fan
// triangles (engine.hpp:1342-1417) or marching pentatope (engine.hpp:1210-1279).
template <class T>
inline void feed(const Params<T>& P, const Mesh& mesh, Accum& acc, int nthreads) {
  const int d = mesh.d;
  const int m = P.m;
  const uint64_t nhc = mesh.n_hypercubes();
  const Chains ch = make_chains(d, m);
  if (ch.n == 0) throw std::runtime_error("unsupported (d, m)");
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  std::vector<std::vector<Node>> tl_nodes(nthreads);
  std::vector<std::array<std::vector<std::vector<Simplex>>, 4>> tl_man(nthreads);

  // 1. LUT pass.
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t hc = 0; hc < (int64_t)nhc; ++hc) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    uint64_t b[4];
    mesh.hc_base((uint64_t)hc, b);
    for (int type = 0; type < ch.n; ++type) {
      Simplex s;
      s.dimension = m;
      for (int vi = 0; vi <= m; ++vi) {
        uint64_t g[4] = {0, 0, 0, 0};
        for (int j = 0; j < d; ++j) g[j] = b[j] + mesh.off[j] + ((ch.s[type][vi] >> j) & 1);
        s.v[vi] = mesh.id(g);
      }
      s.sort_vertices();
      Node nd;
      if (extract(P, mesh, s, nd)) tl_nodes[tid].push_back(nd);
    }
  }

  // 2. Manifold pass over all d! cells of all hypercubes.
  int nperm = 1;
  for (int i = 2; i <= d; ++i) nperm *= i;
  const int k = d - m;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t hc = 0; hc < (int64_t)nhc; ++hc) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    auto& nodes_out = tl_nodes[tid];
    auto& man = tl_man[tid];
    uint64_t b[4];
    mesh.hc_base((uint64_t)hc, b);
    for (int p_idx = 0; p_idx < nperm; ++p_idx) {
      Simplex cell = d_cell(mesh, b, p_idx);
      if (m == 1 && (d == 4 || d == 3)) {
        // marching pentatope / tetrahedron: contour only; any other predicate
        // returns early (engine.hpp:1212-1214, 1283).
        if (P.kind != PRED_CONTOUR) continue;
        const int nv = d + 1;
        T vals[5];
        std::vector<int> A, B;
        for (int i = 0; i < nv; ++i) {
          uint64_t g[4], loc[4];
          mesh.coords(cell.v[i], g);
          for (int j = 0; j < 4; ++j) loc[j] = g[j] - (j < d ? mesh.off[j] : 0);
          vals[i] = P.comp[0].at(loc, d) - P.thr[0];
          if (sos::sign(vals[i], cell.v[i], P.Q) > 0)
            A.push_back(i);
          else
            B.push_back(i);
        }
        auto edge = [&](int a, int c) {
          Simplex e;
          e.dimension = 1;
          e.v[0] = std::min(cell.v[a], cell.v[c]);
          e.v[1] = std::max(cell.v[a], cell.v[c]);
          return e;
        };
        auto ensure = [&](const Simplex& e, std::vector<Simplex>& list) {
          Node nd;
          if (extract(P, mesh, e, nd)) {
            nodes_out.push_back(nd);
            list.push_back(e);
          }
        };
        if (d == 4) {
          if (A.size() == 1 || B.size() == 1) {
            int single = (A.size() == 1) ? A[0] : B[0];
            const auto& others = (A.size() == 1) ? B : A;
            std::vector<Simplex> tet;
            for (int o : others) ensure(edge(single, o), tet);
            if (tet.size() == 4) man[3].push_back(tet);
          } else if (A.size() == 2 || B.size() == 2) {
            const auto& two = (A.size() == 2) ? A : B;
            const auto& three = (A.size() == 2) ? B : A;
            std::vector<Simplex> T0, T1;
            for (int t : three) {
              ensure(edge(two[0], t), T0);
              ensure(edge(two[1], t), T1);
            }
            if (T0.size() == 3 && T1.size() == 3) {
              man[3].push_back({T0[0], T0[1], T0[2], T1[2]});
              man[3].push_back({T0[0], T0[1], T1[1], T1[2]});
              man[3].push_back({T0[0], T1[0], T1[1], T1[2]});
            }
          }
        } else {  // d == 3, marching_tetrahedron (engine.hpp:1281-1340)
          if (A.size() == 1 || B.size() == 1) {
            int single = (A.size() == 1) ? A[0] : B[0];
            const auto& others = (A.size() == 1) ? B : A;
            std::vector<Simplex> tri;
            for (int o : others) ensure(edge(single, o), tri);
            if (tri.size() == 3) man[2].push_back(tri);
          } else if (A.size() == 2 && B.size() == 2) {
            // The reference orders the four nodes by discovery-order node id; the
            // restatement uses Simplex order (equal whenever the four were found by
            // the LUT pass of one feed, which assigns ids in Simplex order).
            std::vector<Simplex> q;
            for (int a : A)
              for (int c : B) ensure(edge(a, c), q);
            if (q.size() == 4) {
              std::sort(q.begin(), q.end());
              man[2].push_back({q[0], q[1], q[2]});
              man[2].push_back({q[0], q[2], q[3]});
            }
          }
        }
        continue;
      }
      // form_general_manifold_patches (engine.hpp:1342-1417)
      if (k <= 0) continue;
      struct NI {
        uint64_t code;
        Simplex s;
      };
      std::vector<NI> ns;
      // find_m_subsimplices (engine.hpp:1645-1652): lexicographic index combos
      const int n = d + 1, r = m + 1;
      int c[5];
      for (int i = 0; i < r; ++i) c[i] = i;
      while (c[0] <= n - r) {
        Simplex f;
        f.dimension = m;
        for (int i = 0; i < r; ++i) f.v[i] = cell.v[c[i]];
        f.sort_vertices();
        Node nd;
        if (extract(P, mesh, f, nd)) {
          nodes_out.push_back(nd);
          ns.push_back({encode_simplex_id(f, mesh), f});
        }
        int i = r - 1;
        while (i >= 0 && c[i] == n - r + i) i--;
        if (i < 0) break;
        c[i]++;
        for (int j = i + 1; j < r; ++j) c[j] = c[i] + j - i;
      }
      if (ns.empty()) continue;
      std::stable_sort(ns.begin(), ns.end(), [](const NI& a, const NI& b2) { return a.code < b2.code; });
      const Simplex hS = ns[0].s;
      if (k == 1) {
        for (size_t i = 1; i < ns.size(); ++i) man[1].push_back({hS, ns[i].s});
      } else if (k == 2) {
        for (int i = 0; i <= d; ++i) {
          std::vector<Simplex> nT;
          for (const auto& x : ns) {
            bool on = true;
            for (int j = 0; j <= m && on; ++j) {
              bool found = false;
              for (int l = 0; l <= d; ++l)
                if (l != i && x.s.v[j] == cell.v[l]) {
                  found = true;
                  break;
                }
              if (!found) on = false;
            }
            if (on) nT.push_back(x.s);
          }
          if (nT.size() < 2) continue;
          const Simplex hT = nT[0];
          for (size_t j = 1; j < nT.size(); ++j) {
            if (!(hS == hT) && !(hS == nT[j])) man[2].push_back({hS, hT, nT[j]});
          }
        }
      }
    }
  }

  for (int t = 0; t < nthreads; ++t) {
    acc.nodes.insert(acc.nodes.end(), tl_nodes[t].begin(), tl_nodes[t].end());
    std::vector<Node>().swap(tl_nodes[t]);
    for (int dm = 1; dm <= 3; ++dm)
      acc.manifold[dm].insert(acc.manifold[dm].end(), tl_man[t][dm].begin(), tl_man[t][dm].end());
  }
  // Nodes are unique by Simplex (engine.hpp:149-157 active_nodes_ map).
  std::sort(acc.nodes.begin(), acc.nodes.end());
  acc.nodes.erase(std::unique(acc.nodes.begin(), acc.nodes.end(),
                              [](const Node& a, const Node& b2) { return a.s == b2.s; }),
                  acc.nodes.end());
}

// UnionFind (engine.hpp:46-74): unite() hangs the larger root under the smaller,
// so every root is the minimum id of its set.
struct UnionFind {
  std::vector<uint64_t> parent;
  void reset(size_t n) {
    parent.resize(n);
    std::iota(parent.begin(), parent.end(), 0);
  }
  uint64_t find(uint64_t i) {
    uint64_t r = i;
    while (parent[r] != r) r = parent[r];
    while (parent[i] != r) {
      uint64_t nx = parent[i];
      parent[i] = r;
      i = nx;
    }
    return r;
  }
  void unite(uint64_t i, uint64_t j) {
    uint64_t a = find(i), b = find(j);
    if (a != b) {
      if (a < b)
        parent[b] = a;
      else
        parent[a] = b;
    }
  }
};

struct Complex {
  std::vector<Node> vertices;  // sorted by Simplex
  std::vector<uint64_t> track_id;
  struct Conn {
    int dimension;
    std::vector<uint32_t> indices;
  };
  std::vector<Conn> connectivity;  // sorted by dimension
  uint64_t n_manifold_unions = 0;
  uint64_t n_components = 0;
};

// get_complex (engine.hpp:822-996) + FeatureComplex::sort (complex.hpp:27-58).
inline Complex get_complex(Accum& acc) {
  Complex cx;
  cx.vertices = acc.nodes;  // already sorted & unique by Simplex
  const size_t n = cx.vertices.size();
  auto index_of = [&](const Simplex& s) -> uint32_t {
    auto it = std::lower_bound(cx.vertices.begin(), cx.vertices.end(), s,
                               [](const Node& a, const Simplex& b2) { return a.s < b2; });
    if (it == cx.vertices.end() || !(it->s == s)) throw std::runtime_error("manifold node not active");
    return (uint32_t)(it - cx.vertices.begin());
  };
  UnionFind uf;
  uf.reset(n);
  for (int dm = 1; dm <= 3; ++dm) {
    std::vector<std::vector<uint32_t>> cells;
    cells.reserve(acc.manifold[dm].size());
    for (const auto& c : acc.manifold[dm]) {
      std::vector<uint32_t> idx;
      for (const auto& s : c) idx.push_back(index_of(s));
      std::sort(idx.begin(), idx.end());
      cells.push_back(idx);
    }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    for (const auto& c : cells)
      for (size_t i = 1; i < c.size(); ++i) {
        uf.unite(c[0], c[i]);
        cx.n_manifold_unions++;
      }
    if (!cells.empty()) {
      Complex::Conn conn;
      conn.dimension = dm;
      for (const auto& c : cells) conn.indices.insert(conn.indices.end(), c.begin(), c.end());
      cx.connectivity.push_back(std::move(conn));
    }
  }
  cx.track_id.resize(n);
  for (size_t i = 0; i < n; ++i) {
    cx.track_id[i] = uf.find(i);
    if (cx.track_id[i] == i) cx.n_components++;
  }
  return cx;
}

// execute (engine.hpp:96-99).
template <class T>
inline Complex execute(const Params<T>& P, const Mesh& mesh, int nthreads) {
  Accum acc;
  if (mesh.has_hypercubes()) feed(P, mesh, acc, nthreads);
  return get_complex(acc);
}

// execute_stream (engine.hpp:214-270): per [t, t+1] slab, copy prev|curr into a
// slab array of shape spatial+[2] and feed a slab mesh with offset t in time.
// `steps[t][c]` points at timestep t of component c (spatial, x fastest, contiguous).
template <class T>
inline Complex execute_stream(Params<T> P, int sd, const uint64_t* spatial, int n_t,
                              const std::vector<std::vector<const T*>>& steps,
                              const std::vector<const T*>& scalar_steps, int nthreads) {
  Accum acc;
  if (n_t < 2) return get_complex(acc);
  uint64_t nsp = 1;
  for (int j = 0; j < sd; ++j) nsp *= spatial[j];
  std::vector<std::vector<T>> slab(P.ncomp + 1, std::vector<T>(2 * nsp));
  for (int t = 0; t + 1 < n_t; ++t) {
    Mesh mesh;
    mesh.d = sd + 1;
    for (int j = 0; j < sd; ++j) {
      mesh.dims[j] = spatial[j];
      mesh.gdims[j] = spatial[j];
      mesh.off[j] = 0;
    }
    mesh.dims[sd] = 2;
    mesh.off[sd] = (uint64_t)t;
    mesh.gdims[sd] = (uint64_t)n_t;
    Params<T> Q = P;
    for (int c = 0; c < P.ncomp; ++c) {
      std::copy(steps[t][c], steps[t][c] + nsp, slab[c].begin());
      std::copy(steps[t + 1][c], steps[t + 1][c] + nsp, slab[c].begin() + nsp);
      Q.comp[c].p = slab[c].data();
      int64_t st = 1;
      for (int j = 0; j < sd; ++j) {
        Q.comp[c].st[j] = st;
        st *= (int64_t)spatial[j];
      }
      Q.comp[c].st[sd] = st;
    }
    Q.has_scalar = false;
    (void)scalar_steps;
    feed(Q, mesh, acc, nthreads);
  }
  return get_complex(acc);
}

}  // namespace oracle
