// ORACLE — TEST INFRASTRUCTURE ONLY.  C entry points over oracle.hpp so that the
// Python tests (ctypes) and bench.py's cpu_baseline leg can drive the CPU
// restatement.  Never linked into the product library.
#include <cstring>
#include <memory>
#include <string>

#include "lean.hpp"
#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;

struct Handle {
  Complex cx;
  int m = 0;
};

template <class T>
Params<T> make_params(int kind, int m, double Q, const double* thr, int ncomp, const void* const* ptrs,
                      const int64_t* strides, const void* sptr, const int64_t* sstrides) {
  Params<T> P;
  P.kind = kind;
  P.m = m;
  P.Q = Q;
  if (thr) {
    P.thr[0] = (T)thr[0];
    P.thr[1] = (T)thr[1];
  }
  P.ncomp = ncomp;
  for (int c = 0; c < ncomp && c < 3; ++c) {
    P.comp[c].p = static_cast<const T*>(ptrs[c]);
    for (int j = 0; j < 4; ++j) P.comp[c].st[j] = strides[c * 4 + j];
  }
  if (sptr) {
    P.has_scalar = true;
    P.scalar.p = static_cast<const T*>(sptr);
    for (int j = 0; j < 4; ++j) P.scalar.st[j] = sstrides[j];
  }
  return P;
}

template <class T>
void load(T (&dst)[5][4], const double* src, int rows, int cols) {
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) dst[i][j] = (T)src[i * cols + j];
}
}  // namespace

extern "C" {

const char* oracle_error() { return g_err.c_str(); }

void* oracle_run(int d, const uint64_t* dims, const uint64_t* off, const uint64_t* gdims, int kind, int m,
                 int dtype, double Q, const double* thr, int ncomp, const void* const* ptrs,
                 const int64_t* strides, const void* sptr, const int64_t* sstrides, int nthreads,
                 uint64_t tstride) {
  try {
    Mesh mesh;
    mesh.d = d;
    mesh.tstride = tstride;
    for (int j = 0; j < d; ++j) {
      mesh.dims[j] = dims[j];
      mesh.off[j] = off ? off[j] : 0;
      mesh.gdims[j] = gdims ? gdims[j] : dims[j];
    }
    auto h = std::make_unique<Handle>();
    h->m = m;
    if (dtype == 0)
      h->cx = execute(make_params<float>(kind, m, Q, thr, ncomp, ptrs, strides, sptr, sstrides), mesh, nthreads);
    else
      h->cx = execute(make_params<double>(kind, m, Q, thr, ncomp, ptrs, strides, sptr, sstrides), mesh, nthreads);
    return h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// ptrs[t * ncomp + c]: timestep t, component c; each a contiguous spatial array.
void* oracle_run_stream(int sd, const uint64_t* spatial, int nt, int kind, int m, int dtype, double Q,
                        const double* thr, int ncomp, const void* const* ptrs, int nthreads) {
  try {
    auto h = std::make_unique<Handle>();
    h->m = m;
    auto run = [&](auto tag) {
      using T = decltype(tag);
      Params<T> P;
      P.kind = kind;
      P.m = m;
      P.Q = Q;
      if (thr) {
        P.thr[0] = (T)thr[0];
        P.thr[1] = (T)thr[1];
      }
      P.ncomp = ncomp;
      std::vector<std::vector<const T*>> steps(nt, std::vector<const T*>(ncomp));
      for (int t = 0; t < nt; ++t)
        for (int c = 0; c < ncomp; ++c) steps[t][c] = static_cast<const T*>(ptrs[t * ncomp + c]);
      h->cx = execute_stream(P, sd, spatial, nt, steps, {}, nthreads);
    };
    if (dtype == 0)
      run(float{});
    else
      run(double{});
    return h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// Full-size mode (lean.hpp): same arguments as oracle_run, star-edge predicates only.
void* oracle_lean_run(int d, const uint64_t* dims, const uint64_t* off, const uint64_t* gdims, int kind, int m,
                      int dtype, double Q, const double* thr, int ncomp, const void* const* ptrs,
                      const int64_t* strides, int nthreads) {
  try {
    Mesh mesh;
    mesh.d = d;
    for (int j = 0; j < d; ++j) {
      mesh.dims[j] = dims[j];
      mesh.off[j] = off ? off[j] : 0;
      mesh.gdims[j] = gdims ? gdims[j] : dims[j];
    }
    auto L = std::make_unique<Lean>();
    if (dtype == 0)
      execute_lean(make_params<float>(kind, m, Q, thr, ncomp, ptrs, strides, nullptr, nullptr), mesh, nthreads, *L);
    else
      execute_lean(make_params<double>(kind, m, Q, thr, ncomp, ptrs, strides, nullptr, nullptr), mesh, nthreads, *L);
    return L.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_lean_sizes(void* hp, uint64_t* n_nodes, uint64_t* n_edges, uint64_t* n_components) {
  const auto* L = static_cast<const Lean*>(hp);
  *n_nodes = L->keys.size();
  *n_edges = L->edges.size();
  *n_components = L->n_components;
}

// Nodes [begin, begin + count): sorted vertex ids, lambda, track ids (any pointer may be NULL).
void oracle_lean_nodes(void* hp, uint64_t begin, uint64_t count, uint64_t* verts, double* lam, uint64_t* track) {
  const auto* L = static_cast<const Lean*>(hp);
  const int nv = L->m + 1;
  for (uint64_t i = 0; i < count; ++i) {
    if (verts) L->vertices(begin + i, verts + i * nv);
    for (int k = 0; k < nv && lam; ++k) lam[i * nv + k] = L->lambda[(begin + i) * nv + k];
    if (track) track[i] = L->track[begin + i];
  }
}

// Edges [begin, begin + count) as (a, b) uint32 pairs.
void oracle_lean_edges(void* hp, uint64_t begin, uint64_t count, uint32_t* out) {
  const auto* L = static_cast<const Lean*>(hp);
  for (uint64_t i = 0; i < count; ++i) {
    out[2 * i] = (uint32_t)(L->edges[begin + i] >> 32);
    out[2 * i + 1] = (uint32_t)L->edges[begin + i];
  }
}

void oracle_lean_free(void* hp) { delete static_cast<Lean*>(hp); }

uint64_t oracle_n_nodes(void* hp) { return static_cast<Handle*>(hp)->cx.vertices.size(); }
int oracle_node_dim(void* hp) { return static_cast<Handle*>(hp)->m; }

void oracle_nodes(void* hp, uint64_t* verts, double* lam, float* bary, float* scalar, uint64_t* track) {
  auto* h = static_cast<Handle*>(hp);
  const int nv = h->m + 1;
  for (size_t i = 0; i < h->cx.vertices.size(); ++i) {
    const Node& n = h->cx.vertices[i];
    for (int k = 0; k < nv; ++k) {
      if (verts) verts[i * nv + k] = n.s.v[k];
      if (lam) lam[i * nv + k] = n.lambda[k];
      if (bary) bary[i * nv + k] = n.bary[k];
    }
    if (scalar) scalar[i] = n.scalar;
    if (track) track[i] = h->cx.track_id[i];
  }
}

int oracle_n_conn(void* hp) { return (int)static_cast<Handle*>(hp)->cx.connectivity.size(); }
int oracle_conn_dim(void* hp, int i) { return static_cast<Handle*>(hp)->cx.connectivity[i].dimension; }
uint64_t oracle_conn_size(void* hp, int i) { return static_cast<Handle*>(hp)->cx.connectivity[i].indices.size(); }
void oracle_conn(void* hp, int i, uint32_t* out) {
  const auto& v = static_cast<Handle*>(hp)->cx.connectivity[i].indices;
  std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
}
uint64_t oracle_n_components(void* hp) { return static_cast<Handle*>(hp)->cx.n_components; }
void oracle_free(void* hp) { delete static_cast<Handle*>(hp); }

// ---- unit-level entry points (values arrive as doubles; dtype 0 narrows to float)
int oracle_sign(int dtype, double v, uint64_t idx, double Q) {
  return dtype == 0 ? sos::sign<float>((float)v, idx, Q) : sos::sign<double>(v, idx, Q);
}

int oracle_det2(int dtype, const double* a, const double* b, uint64_t i0, uint64_t i1, double Q) {
  if (dtype == 0) {
    float x[2] = {(float)a[0], (float)a[1]}, y[2] = {(float)b[0], (float)b[1]};
    return sos::det2(x, y, i0, i1, Q);
  }
  return sos::det2(a, b, i0, i1, Q);
}

int oracle_det3(int dtype, const double* a, const double* b, const double* c, double Q) {
  if (dtype == 0) {
    float x[3], y[3], z[3];
    for (int j = 0; j < 3; ++j) {
      x[j] = (float)a[j];
      y[j] = (float)b[j];
      z[j] = (float)c[j];
    }
    return sos::det3(x, y, z, Q);
  }
  return sos::det3(a, b, c, Q);
}

// Level-0 value of det3 (the 4x4 determinant) as hi:lo 64-bit halves.
void oracle_det3_level0(const int64_t* qa, const int64_t* qb, const int64_t* qc, int64_t* hi, uint64_t* lo) {
  i128 M[16];
  const int64_t* r[3] = {qa, qb, qc};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M[i * 4 + j] = r[i][j];
    M[i * 4 + 3] = 1;
  }
  M[12] = M[13] = M[14] = 0;
  M[15] = 1;
  i128 v = sos::idet(M, 4);
  *hi = (int64_t)(v >> 64);
  *lo = (uint64_t)v;
}

int oracle_origin_inside(int K, int dtype, const double* vals, const uint64_t* idx, double Q) {
  auto go = [&](auto tag) -> int {
    using T = decltype(tag);
    T v[5][4];
    load(v, vals, K + 1, K);
    if (K == 1) {
      T w[2][1] = {{v[0][0]}, {v[1][0]}};
      return sos::origin_inside<1, T>::check(w, idx, Q);
    }
    if (K == 2) {
      T w[3][2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j) w[i][j] = v[i][j];
      return sos::origin_inside<2, T>::check(w, idx, Q);
    }
    if (K == 3) {
      T w[4][3];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 3; ++j) w[i][j] = v[i][j];
      return sos::origin_inside<3, T>::check(w, idx, Q);
    }
    return -1;
  };
  return dtype == 0 ? go(float{}) : go(double{});
}

// origin_inside<K> on n simplices at once: vals[n][K+1][K], idx[n][K+1] -> out[n].
void oracle_origin_inside_many(int K, int dtype, int64_t n, const double* vals, const uint64_t* idx, double Q,
                               int* out) {
  for (int64_t s = 0; s < n; ++s) out[s] = oracle_origin_inside(K, dtype, vals + s * (K + 1) * K, idx + s * (K + 1), Q);
}

// Returns 1 if the solver succeeds; lambda (K+1 entries) is written either way.
int oracle_solve(int K, int dtype, const double* vals, double* lambda) {
  auto go = [&](auto tag) -> int {
    using T = decltype(tag);
    T v[5][4];
    load(v, vals, K + 1, K);
    T lam[5] = {0, 0, 0, 0, 0};
    int ok = -1;
    if (K == 1) {
      T w[2][1] = {{v[0][0]}, {v[1][0]}};
      ok = Solver<1, T>::solve(w, lam);
    } else if (K == 2) {
      T w[3][2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j) w[i][j] = v[i][j];
      ok = Solver<2, T>::solve(w, lam);
    } else if (K == 3) {
      T w[4][3];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 3; ++j) w[i][j] = v[i][j];
      ok = Solver<3, T>::solve(w, lam);
    }
    for (int i = 0; i <= K; ++i) lambda[i] = (double)lam[i];
    return ok;
  };
  return dtype == 0 ? go(float{}) : go(double{});
}

int oracle_chains(int d, int m, uint8_t* out) {
  Chains c = make_chains(d, m);
  for (int t = 0; t < c.n; ++t)
    for (int i = 0; i <= m; ++i) out[t * (m + 1) + i] = c.s[t][i];
  return c.n;
}

uint64_t oracle_encode_simplex_id(int d, const uint64_t* gdims, int dim, const uint64_t* verts) {
  Mesh mesh;
  mesh.d = d;
  for (int j = 0; j < d; ++j) mesh.gdims[j] = mesh.dims[j] = gdims[j];
  Simplex s;
  s.dimension = dim;
  for (int i = 0; i <= dim; ++i) s.v[i] = verts[i];
  return encode_simplex_id(s, mesh);
}
}
