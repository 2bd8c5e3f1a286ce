// Compile-time tables of the Freudenthal (Kuhn) triangulation of a D-dimensional
// regular grid, D <= 4.
//
//  * Simplex types.  An m-simplex whose smallest vertex is b has vertices
//    b + s0, ..., b + sm with 0 = s0 < s1 < ... < sm a chain of axis sets
//    ({0,1}^D bit masks, bit j = axis j).  These are the rows of the reference
//    LUTs cpu_lut_D_M (include/ftk2/core/mesh.hpp:111-229); the counts are
//    3,2,7,12,6,15,50,60,24 (mesh.hpp:232-236).  Types are numbered in
//    lexicographic order of (s1, ..., sm) read as binary numbers, which is the
//    order of Simplex::operator< (mesh.hpp:36-42) among simplices with the same
//    smallest vertex: vertex ids grow with the binary value of the offset
//    whenever every global extent is >= 2.  So (v_min, type) is the canonical
//    sort key of a node.
//  * Kuhn cells.  D! cells per hypercube, cell p = chain 0 = c0 < c1 < ... < cD =
//    all axes, the permutation decoded like get_d_simplex (mesh.hpp:403-414).
//  * Faces of a cell.  The (M+1)-subsets of cell vertices in the order of
//    find_m_subsimplices (engine.hpp:1645-1652).  For each (cell, face): the
//    owner offset c_{i0} (smallest vertex), the type of the face at its owner,
//    and its rank in encode_simplex_id order (mesh.hpp:736-749), which decides
//    the star centre of form_general_manifold_patches (engine.hpp:1376-1382).
#pragma once
#include <cstdint>

namespace ftk2h {

constexpr int factorial(int n) {
  int r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
constexpr int binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

template <int D, int M>
struct ChainTable {
  int n = 0;
  uint8_t s[64][M + 1] = {};
};

template <int D, int M>
constexpr ChainTable<D, M> build_chains() {
  ChainTable<D, M> t{};
  const int full = (1 << D) - 1;
  int idx[M + 1] = {};
  for (int i = 1; i <= M; ++i) idx[i] = 1;
  while (true) {
    bool ok = true;
    for (int i = 1; i <= M && ok; ++i) {
      const int prev = idx[i - 1], cur = idx[i];
      if ((cur & prev) != prev || cur == prev) ok = false;
    }
    if (ok) {
      for (int i = 0; i <= M; ++i) t.s[t.n][i] = (uint8_t)idx[i];
      t.n++;
    }
    int p = M;
    while (p >= 1) {
      if (idx[p] < full) {
        idx[p]++;
        break;
      }
      idx[p] = 1;
      p--;
    }
    if (p < 1) break;
  }
  return t;
}

template <int D, int M>
struct Chains {
  static constexpr ChainTable<D, M> table = build_chains<D, M>();
  static constexpr int N = table.n;
};

template <int D, int M>
constexpr int chain_index_in(const ChainTable<D, M>& T, const uint8_t* rel) {
  for (int t = 0; t < T.n; ++t) {
    bool eq = true;
    for (int i = 0; i <= M; ++i)
      if (T.s[t][i] != rel[i]) eq = false;
    if (eq) return t;
  }
  return -1;
}

template <int D, int M>
struct CellFaceTable {
  static constexpr int NP = factorial(D);
  static constexpr int NF = binom(D + 1, M + 1);
  uint8_t cell[NP][D + 1] = {};     // chain offsets of the cell vertices
  uint8_t pos[NF][M + 1] = {};      // face -> cell vertex positions
  uint8_t vmask[NF] = {};           // face -> bit mask of cell vertex positions
  uint8_t owner[NP][NF] = {};       // offset of the face's smallest vertex
  uint8_t type[NP][NF] = {};        // face type at its owner
  uint8_t order[NP][NF] = {};       // faces sorted by encode_simplex_id
};

template <int D, int M>
constexpr CellFaceTable<D, M> build_cell_faces() {
  CellFaceTable<D, M> t{};
  const ChainTable<D, M> chains = build_chains<D, M>();
  constexpr int NP = CellFaceTable<D, M>::NP;
  constexpr int NF = CellFaceTable<D, M>::NF;
  for (int p_idx = 0; p_idx < NP; ++p_idx) {
    int p[4] = {0, 1, 2, 3};
    int tmp = p_idx;
    for (int i = 0; i < D; ++i) {
      const int j = tmp % (D - i);
      const int sw = p[i];
      p[i] = p[i + j];
      p[i + j] = sw;
      tmp /= (D - i);
    }
    t.cell[p_idx][0] = 0;
    for (int i = 0; i < D; ++i) t.cell[p_idx][i + 1] = (uint8_t)(t.cell[p_idx][i] | (1 << p[i]));
  }
  // faces: lexicographic (M+1)-combinations of D+1 positions
  int c[5] = {0, 1, 2, 3, 4};
  const int n = D + 1, r = M + 1;
  int f = 0;
  while (true) {
    uint8_t vm = 0;
    for (int i = 0; i < r; ++i) {
      t.pos[f][i] = (uint8_t)c[i];
      vm = (uint8_t)(vm | (1 << c[i]));
    }
    t.vmask[f] = vm;
    f++;
    int i = r - 1;
    while (i >= 0 && c[i] == n - r + i) i--;
    if (i < 0) break;
    c[i]++;
    for (int j = i + 1; j < r; ++j) c[j] = c[i] + j - i;
  }
  for (int p_idx = 0; p_idx < NP; ++p_idx) {
    uint32_t code[NF] = {};
    for (int fi = 0; fi < NF; ++fi) {
      const uint8_t own = t.cell[p_idx][t.pos[fi][0]];
      uint8_t rel[M + 1] = {};
      uint32_t comb = 0;
      for (int i = 0; i <= M; ++i) {
        rel[i] = (uint8_t)(t.cell[p_idx][t.pos[fi][i]] ^ own);
        if (i > 0) comb = (comb << 4) | rel[i];
      }
      t.owner[p_idx][fi] = own;
      t.type[p_idx][fi] = (uint8_t)chain_index_in<D, M>(chains, rel);
      code[fi] = ((uint32_t)own << 24) | (comb & 0xFFFFFF);
    }
    // order faces by code (codes are distinct within a cell)
    bool used[NF] = {};
    for (int k = 0; k < NF; ++k) {
      int best = -1;
      for (int fi = 0; fi < NF; ++fi)
        if (!used[fi] && (best < 0 || code[fi] < code[best])) best = fi;
      used[best] = true;
      t.order[p_idx][k] = (uint8_t)best;
    }
  }
  return t;
}

template <int D, int M>
struct CellFaces {
  static constexpr CellFaceTable<D, M> table = build_cell_faces<D, M>();
  static constexpr int NP = CellFaceTable<D, M>::NP;
  static constexpr int NF = CellFaceTable<D, M>::NF;
};

// Number of chain types for runtime (d, m) checks.
inline int n_types_rt(int d, int m) {
  static const int tab[5][5] = {{0, 0, 0, 0, 0},
                                {0, 1, 0, 0, 0},
                                {0, 3, 2, 0, 0},
                                {0, 7, 12, 6, 0},
                                {0, 15, 50, 60, 24}};
  if (d < 1 || d > 4 || m < 1 || m > d) return 0;
  return tab[d][m];
}

}  // namespace ftk2h
