// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Full-size mode of the oracle for predicates whose manifold cells are star edges
// (k = d - m = 1: CriticalPointPredicate<2> on 2-D + time, CriticalPointPredicate<3>
// on 3-D + time, FiberPredicate on 2-D + time).  It computes the same FeatureComplex
// as oracle::execute() (oracle.hpp) with the same predicate evaluation (extract(),
// i.e. sos.hpp origin_inside + zero_crossing.hpp solver + predicate.hpp extract_it),
// the same star-edge rule (engine.hpp:1342-1387 form_general_manifold_patches with
// encode_simplex_id, mesh.hpp:736-749) and the same union-find (engine.hpp:46-74),
// but with a compact layout so that ~10^8 nodes fit in host memory:
//   * every simplex is evaluated once, by the vertex that owns it (its smallest
//     vertex); the node set is every m-face of every d-cell that passes, which for
//     a mesh with at least two vertices per axis is every chain simplex whose top
//     vertex lies in the grid (oracle.hpp feed() reaches the same set through its
//     LUT pass plus its manifold pass, engine.hpp:1347-1367);
//   * a node is (owner vertex, rank of its type in Simplex order); per-vertex bit
//     masks over ranks and an exclusive prefix of their popcounts give canonical
//     indices (FeatureComplex::vertices order, complex.hpp:27-58) without a map;
//   * cells are (a, b) index pairs; each pair of distinct tets of a pentatope (or
//     triangles of a tet) lies in exactly one d-cell, so no duplicate arises, and
//     the pairs are sorted as get_complex() sorts them (engine.hpp:822-996).
// tests/test_oracle_lean.py checks lean == execute() on small meshes (zero and
// duplicated vertices, offsets, FP32/FP64, fiber) before this mode is trusted at
// full size.
#pragma once
#include <parallel/algorithm>

#include "oracle.hpp"

namespace oracle {

struct Lean {
  int d = 0, m = 0, ntypes = 0;
  Mesh mesh;
  std::vector<uint64_t> keys;   // local vertex index * 64 + rank, ascending
  std::vector<double> lambda;   // (m+1) per node, the solver output widened to double
  std::vector<uint32_t> track;  // FeatureElement::track_id (component minimum)
  std::vector<uint64_t> edges;  // (a << 32) | b, a < b, ascending
  int rank_to_type[64];
  Chains ch;
  uint64_t n_components = 0;

  // Sorted global vertex ids of node i.
  void vertices(uint64_t i, uint64_t* out) const {
    const uint64_t lv = keys[i] >> 6;
    const int t = rank_to_type[keys[i] & 63];
    uint64_t l[4] = {0, 0, 0, 0}, r = lv;
    for (int j = 0; j < d; ++j) {
      l[j] = r % mesh.dims[j];
      r /= mesh.dims[j];
    }
    for (int k = 0; k <= m; ++k) {
      uint64_t g[4] = {0, 0, 0, 0};
      for (int j = 0; j < d; ++j) g[j] = l[j] + mesh.off[j] + ((ch.s[t][k] >> j) & 1);
      out[k] = mesh.id(g);
    }
  }
};

template <class T>
inline void execute_lean(const Params<T>& P, const Mesh& mesh, int nthreads, Lean& L) {
  const int d = mesh.d, m = P.m;
  if (d - m != 1 || P.kind == PRED_CONTOUR) throw std::runtime_error("lean mode: star-edge predicates only");
  if (!mesh.has_hypercubes()) throw std::runtime_error("lean mode: mesh without hypercubes");
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  L.d = d;
  L.m = m;
  L.mesh = mesh;
  L.ch = make_chains(d, m);
  const Chains& ch = L.ch;
  L.ntypes = ch.n;
  // Simplex order of the types of one vertex: vertex ids are v + id offset, so the
  // order does not depend on v; take it from vertex 0 of the mesh.
  int rank[64];
  {
    std::vector<std::pair<Simplex, int>> v;
    for (int t = 0; t < ch.n; ++t) {
      Simplex s;
      s.dimension = m;
      for (int k = 0; k <= m; ++k) {
        uint64_t g[4] = {0, 0, 0, 0};
        for (int j = 0; j < d; ++j) g[j] = mesh.off[j] + ((ch.s[t][k] >> j) & 1);
        s.v[k] = mesh.id(g);
      }
      s.sort_vertices();
      v.push_back({s, t});
    }
    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (int r = 0; r < ch.n; ++r) {
      rank[v[r].second] = r;
      L.rank_to_type[r] = v[r].second;
    }
  }
  // chain (s1..sm packed 4 bits each) -> type
  std::vector<int> chain_type(1 << (4 * m), -1);
  for (int t = 0; t < ch.n; ++t) {
    uint32_t key = 0;
    for (int k = 1; k <= m; ++k) key |= (uint32_t)ch.s[t][k] << (4 * (k - 1));
    chain_type[key] = t;
  }

  // 1. nodes: every in-grid chain simplex, evaluated by its owner vertex
  const uint64_t nv = mesh.n_vertices();
  const uint64_t CH = 4096;
  const uint64_t nchunk = (nv + CH - 1) / CH;
  std::vector<uint64_t> mask(nv, 0);
  std::vector<std::vector<double>> lam_chunk(nchunk);
  std::vector<uint64_t> cnt(nchunk + 1, 0);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t c = 0; c < (int64_t)nchunk; ++c) {
    std::vector<double>& out = lam_chunk[c];
    uint64_t n = 0;
    const uint64_t e = std::min(nv, (uint64_t)(c + 1) * CH);
    for (uint64_t v = (uint64_t)c * CH; v < e; ++v) {
      uint64_t l[4] = {0, 0, 0, 0}, r = v;
      for (int j = 0; j < d; ++j) {
        l[j] = r % mesh.dims[j];
        r /= mesh.dims[j];
      }
      Node nd[64];
      bool hit[64] = {};
      uint64_t mk = 0;
      for (int t = 0; t < ch.n; ++t) {
        bool in = true;
        for (int j = 0; j < d; ++j)
          if (((ch.s[t][m] >> j) & 1) && l[j] + 1 >= mesh.dims[j]) in = false;
        if (!in) continue;
        Simplex s;
        s.dimension = m;
        for (int k = 0; k <= m; ++k) {
          uint64_t g[4] = {0, 0, 0, 0};
          for (int j = 0; j < d; ++j) g[j] = l[j] + mesh.off[j] + ((ch.s[t][k] >> j) & 1);
          s.v[k] = mesh.id(g);
        }
        s.sort_vertices();
        if (extract(P, mesh, s, nd[rank[t]])) {
          hit[rank[t]] = true;
          mk |= 1ull << rank[t];
        }
      }
      mask[v] = mk;
      for (int r2 = 0; r2 < ch.n; ++r2)
        if (hit[r2]) {
          for (int k = 0; k <= m; ++k) out.push_back(nd[r2].lambda[k]);
          ++n;
        }
    }
    cnt[c + 1] = n;
  }
  for (uint64_t c = 0; c < nchunk; ++c) cnt[c + 1] += cnt[c];
  const uint64_t nn = cnt[nchunk];
  if (nn >= 0xFFFFFFFFull) throw std::runtime_error("lean mode: more than 2^32-1 nodes");
  L.keys.assign(nn, 0);
  L.lambda.assign(nn * (m + 1), 0.0);
  std::vector<uint32_t> node_off(nv, 0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t c = 0; c < (int64_t)nchunk; ++c) {
    uint64_t pos = cnt[c];
    std::copy(lam_chunk[c].begin(), lam_chunk[c].end(), L.lambda.begin() + pos * (m + 1));
    std::vector<double>().swap(lam_chunk[c]);
    const uint64_t e = std::min(nv, (uint64_t)(c + 1) * CH);
    for (uint64_t v = (uint64_t)c * CH; v < e; ++v) {
      node_off[v] = (uint32_t)pos;
      for (uint64_t mk = mask[v]; mk; mk &= mk - 1) L.keys[pos++] = v * 64 + (uint64_t)__builtin_ctzll(mk);
    }
  }

  // 2. star edges of every d-cell of every hypercube
  const uint64_t nhc = mesh.n_hypercubes();
  int nperm = 1;
  for (int i = 2; i <= d; ++i) nperm *= i;
  std::vector<std::vector<uint64_t>> tl_edges(nthreads);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t hc = 0; hc < (int64_t)nhc; ++hc) {
#ifdef _OPENMP
    auto& out = tl_edges[omp_get_thread_num()];
#else
    auto& out = tl_edges[0];
#endif
    uint64_t b[4];
    mesh.hc_base((uint64_t)hc, b);
    for (int p = 0; p < nperm; ++p) {
      const Simplex cell = d_cell(mesh, b, p);
      struct NI {
        uint64_t code;
        uint32_t idx;
      };
      NI ns[5];
      int cntn = 0;
      // find_m_subsimplices (engine.hpp:1645-1652): drop one cell vertex at a time,
      // in the lexicographic order of the kept index combinations
      for (int drop = d; drop >= 0; --drop) {
        Simplex f;
        f.dimension = m;
        int q = 0;
        for (int i = 0; i <= d; ++i)
          if (i != drop) f.v[q++] = cell.v[i];
        uint64_t g0[4];
        mesh.coords(f.v[0], g0);
        uint32_t key = 0;
        for (int k = 1; k <= m; ++k) {
          uint64_t gk[4];
          mesh.coords(f.v[k], gk);
          uint32_t s = 0;
          for (int j = 0; j < d; ++j) s |= (uint32_t)(gk[j] - g0[j]) << j;
          key |= s << (4 * (k - 1));
        }
        const int t = chain_type[key];
        if (t < 0) throw std::runtime_error("lean mode: face is not a chain simplex");
        uint64_t lv = 0, mul = 1;
        for (int j = 0; j < d; ++j) {
          lv += (g0[j] - mesh.off[j]) * mul;
          mul *= mesh.dims[j];
        }
        const int r = rank[t];
        if (!((mask[lv] >> r) & 1)) continue;
        ns[cntn].code = encode_simplex_id(f, mesh);
        ns[cntn].idx = node_off[lv] + (uint32_t)__builtin_popcountll(mask[lv] & ((1ull << r) - 1));
        ++cntn;
      }
      if (cntn < 2) continue;
      for (int i = 1; i < cntn; ++i)  // stable insertion sort by code
        for (int j = i; j > 0 && ns[j].code < ns[j - 1].code; --j) std::swap(ns[j], ns[j - 1]);
      for (int i = 1; i < cntn; ++i) {
        const uint64_t a = std::min(ns[0].idx, ns[i].idx), c = std::max(ns[0].idx, ns[i].idx);
        out.push_back((a << 32) | c);
      }
    }
  }
  size_t ne = 0;
  for (auto& v : tl_edges) ne += v.size();
  L.edges.clear();
  L.edges.reserve(ne);
  for (auto& v : tl_edges) {
    L.edges.insert(L.edges.end(), v.begin(), v.end());
    std::vector<uint64_t>().swap(v);
  }
  __gnu_parallel::sort(L.edges.begin(), L.edges.end());
  L.edges.erase(std::unique(L.edges.begin(), L.edges.end()), L.edges.end());

  // 3. tracks (engine.hpp:46-74: the larger root goes under the smaller)
  UnionFind uf;
  uf.reset(nn);
  for (uint64_t e : L.edges) uf.unite(e >> 32, e & 0xFFFFFFFFull);
  L.track.assign(nn, 0);
  L.n_components = 0;
  for (uint64_t i = 0; i < nn; ++i) {
    L.track[i] = (uint32_t)uf.find(i);
    if (L.track[i] == i) L.n_components++;
  }
}

}  // namespace oracle
