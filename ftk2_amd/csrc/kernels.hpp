// Device kernels of the simplicial extraction pipeline on a regular spacetime mesh.
//
// One thread per mesh vertex v (local linear index, x fastest).  Vertex v owns every
// simplex whose smallest vertex is v ("node key" = global id(v) * n_types + type),
// and, when v is a hypercube base, the D! Kuhn cells of that hypercube.
//
//   classify   per vertex: bit mask over the n_types simplices owned by v that pass
//              the predicate (exact SoS + solver gate for critical points).  The node
//              set is every m-face of every d-cell that passes: the LUT pass of
//              engine.hpp:118-158 plus the upper-boundary faces that the manifold
//              pass of engine.hpp:1342-1367 adds.
//   offsets    canonical index of each node = exclusive prefix of mask popcounts in
//              vertex order (Simplex order, mesh.hpp:36-42), and per-vertex counts of
//              manifold cells.
//   emit       solver output lambda for every node; manifold cells in canonical
//              indices: star edges / star-fan triangles (engine.hpp:1342-1417) or
//              marching pentatope/tetrahedron tets/triangles (engine.hpp:1210-1340).
//              Every cell of hypercube v contains a node owned by v as its smallest
//              index (k = 1 and marching), so per-vertex segments sorted locally are
//              globally sorted (complex.hpp:27-58); k = 2 cells are sorted globally.
//   link       union-find by hooking the larger root under the smaller
//              (engine.hpp:46-74): root = minimum canonical index = track_id.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "certify.hpp"
#include "exact.hpp"
#include "filter.hpp"
#include "tables.hpp"

namespace ftk2h {

constexpr int KIND_CP = 0, KIND_CONTOUR = 1, KIND_FIBER = 2;
constexpr int BLOCK = 256;
// bsum of a tile that stage 1 leaves to stage 2 (node counts of a tile are below
// 2^14): TILE_PENDING_SOL when its only pending work is the solver gate of SoS-inside
// simplices (stage 1 then writes only pend), TILE_PENDING otherwise.
constexpr uint32_t TILE_PENDING = 0xFFFFFFFFu;
constexpr uint32_t TILE_PENDING_SOL = 0xFFFFFFFEu;

struct KParams {
  int64_t dims[4];     // local extents (1 beyond d)
  int64_t off[4];      // global offset of local (0..0)
  int64_t gstride[4];  // global id strides
  int64_t nvert;       // local vertices
  int small;           // nvert < 2^31: 32-bit index arithmetic
  const void* comp[3];
  int64_t cst[3][4];
  const void* scal;
  int64_t sst[4];
  double Q;
  double thr[2];
  // per-corner offsets (corner s = bit mask of raised axes), precomputed on the host
  int64_t coff[3][16];  // element offset of component c
  int64_t soff[16];     // element offset of the scalar field
  int64_t loff[16];     // local linear vertex offset
  uint64_t goff[16];    // global vertex id offset
  int packed;           // every component laid out x-fastest with local extents
  uint32_t nq_cap;      // emit: queued nodes per workgroup (<= EMIT_NQ)
  uint32_t stage_cap;   // emit: staged cell words per workgroup (<= EMIT_STAGE_WORDS)
  uint64_t* mask;
  uint64_t* pend;   // types left to stage 2 of the classification
  uint64_t* degm;   // those among them that need the verbatim sign levels
  uint16_t* ccount; // manifold cells of the hypercube based at each vertex (k_offsets)
  uint32_t* node_off;
  uint32_t* bsum;   // per-block node counts -> exclusive offsets (+ total at [nb])
  uint32_t* ebsum;  // per-block cell counts  -> exclusive offsets (+ total at [nb])
  uint64_t* keys;
  void* lambda;
  float* scalar_out;
  uint32_t* cells;
  uint32_t* xedges;  // edges still to be united globally: (a, b) pairs
  uint32_t* xcount;  // number of pairs in xedges
  uint32_t* parent;
  uint32_t* track;
  uint64_t n_nodes;
  uint64_t n_cells;
};

// Global-memory view of a device pointer.  Pointers the kernels read out of memory
// (the batched job table) are generic to the compiler, which would then use flat
// instructions; the arrays of KParams always live in global memory.
#define FTK2H_AS1 __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ FTK2H_AS1 T* gptr(T* p) {
  return (FTK2H_AS1 T*)p;
}

// ---------------------------------------------------------------------------
// compile-time loop
template <int I, int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

template <int D>
__device__ __forceinline__ void decompose(const KParams& P, int64_t v, int64_t* l) {
  if (P.small) {
    uint32_t r = (uint32_t)v;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const uint32_t dj = (uint32_t)P.dims[j];
      l[j] = r % dj;
      r /= dj;
    }
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      l[j] = v % P.dims[j];
      v /= P.dims[j];
    }
  }
}

// A vertex as the kernels see it: local linear index and coordinates, element
// index of every field component, global id.  The part a queued job needs is VtxRef.
struct VtxRef {
  int64_t base[3];
  int64_t sbase;
  uint64_t gid;
};
struct Vtx {
  int64_t v;
  int64_t l[4];
  VtxRef r;
};

template <int D>
__device__ __forceinline__ VtxRef make_ref(const KParams& P, int64_t v, const int64_t* l) {
  VtxRef r;
  uint64_t g = 0;
#pragma unroll
  for (int j = 0; j < D; ++j) g += (uint64_t)(l[j] + P.off[j]) * (uint64_t)P.gstride[j];
  r.gid = g;
  if (P.packed) {
    r.base[0] = r.base[1] = r.base[2] = v;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      int64_t e = 0;
#pragma unroll
      for (int j = 0; j < D; ++j) e += l[j] * P.cst[c][j];
      r.base[c] = e;
    }
  }
  int64_t e = 0;
  if (P.scal) {
#pragma unroll
    for (int j = 0; j < D; ++j) e += l[j] * P.sst[j];
  }
  r.sbase = e;
  return r;
}

template <int D>
__device__ __forceinline__ Vtx make_vtx(const KParams& P, int64_t v) {
  Vtx x;
  x.v = v;
  decompose<D>(P, v, x.l);
#pragma unroll
  for (int j = D; j < 4; ++j) x.l[j] = 0;
  x.r = make_ref<D>(P, v, x.l);
  return x;
}

// bit s set <=> corner l + s lies inside the local grid
template <int D>
__device__ __forceinline__ uint32_t corner_mask(const KParams& P, const int64_t* l) {
  uint32_t m = 0;
#pragma unroll
  for (int s = 0; s < (1 << D); ++s) {
    bool in = true;
#pragma unroll
    for (int j = 0; j < D; ++j)
      if ((s >> j) & 1) in = in && (l[j] + 1 < P.dims[j]);
    if (in) m |= 1u << s;
  }
  return m;
}

template <class T>
__device__ __forceinline__ T load_val(const KParams& P, int c, const VtxRef& r, int s) {
  return gptr(static_cast<const T*>(P.comp[c]))[r.base[c] + P.coff[c][s]];
}

template <class T>
__device__ __forceinline__ T load_scalar(const KParams& P, const VtxRef& r, int s) {
  return gptr(static_cast<const T*>(P.scal))[r.sbase + P.soff[s]];
}

// chains packed 4 bits per vertex offset, for run-time type lookups
template <int D, int M>
struct PackedChains {
  uint32_t v[64];
};
template <int D, int M>
constexpr PackedChains<D, M> build_packed() {
  PackedChains<D, M> p{};
  constexpr auto C = build_chains<D, M>();
  for (int t = 0; t < C.n; ++t) {
    uint32_t w = 0;
    for (int i = 1; i <= M; ++i) w |= (uint32_t)C.s[t][i] << (4 * (i - 1));
    p.v[t] = w;
  }
  return p;
}
template <int D, int M>
__device__ __forceinline__ uint32_t packed_chain(int t) {
  constexpr PackedChains<D, M> P = build_packed<D, M>();
  return P.v[t];
}

// Per-corner element offsets and the chain table, copied to LDS once per tile for
// the kernels that index them with per-lane (run-time) corner and type numbers;
// from kernel arguments such indexing would cost a vector load per access.
struct CornerTab {
  int64_t coff[3][16];
  int64_t soff[16];
  int64_t loff[16];
  uint32_t pc[64];
};
template <int D, int M>
__device__ __forceinline__ void load_corner_tab(const KParams& P, CornerTab* tab) {
  for (int i = threadIdx.x; i < 48 + 16 + 16 + 64; i += BLOCK) {
    if (i < 48)
      tab->coff[i / 16][i % 16] = P.coff[i / 16][i % 16];
    else if (i < 64)
      tab->soff[i - 48] = P.soff[i - 48];
    else if (i < 80)
      tab->loff[i - 64] = P.loff[i - 64];
    else
      tab->pc[i - 80] = packed_chain<D, M>(i - 80);
  }
}
template <class T>
__device__ __forceinline__ T load_val_t(const KParams& P, const CornerTab& tab, int c, const VtxRef& r, int s) {
  return gptr(static_cast<const T*>(P.comp[c]))[r.base[c] + tab.coff[c][s]];
}

// Table references usable inside lambdas without capture.
template <int D, int M>
constexpr const ChainTable<D, M>& CH = Chains<D, M>::table;
template <int D, int M>
constexpr const CellFaceTable<D, M>& CF = CellFaces<D, M>::table;

// ---------------------------------------------------------------------------
// Triple product signs: the error-bounded FP32 filter of filter.hpp (cross3f,
// dot_signf); a value within the bound sends its simplex to the exact path.

// Bit-sliced tet logic of classify_cp3.  Triangle (0,a,b) and tet (a,b,c) values
// are computed in type order (types are sorted by their chain, so the tets sharing
// (a,b) are consecutive and follow triangle (a,b)); each value's bit is shifted into
// 32-bit words as it is computed (w = (w << 1) | bit), so item i of n lies in word
// i >> 5 at bit push_bit(i, n).  A tet needs the bits of three triangles, (0,a,b),
// (0,a,c) and (0,b,c) ("slots" 0, 1, 2), at its own bit position.  Every such move
// shifts the 64 triangle bits by d = (tet bit) - (triangle bit); the moves of one
// slot and tet word with the same d form one group, gathered by one shift and one
// masked OR.  The uncertain bits of all three slots are ORed together, so their
// groups are merged across slots.
constexpr int push_bit(int i, int n) {
  const int w = i >> 5;
  const int in_word = n - 32 * w < 32 ? n - 32 * w : 32;
  return in_word - 1 - (i & 31);
}

template <int D>
struct SliceTab {
  int nsg[3][2] = {};           // sign groups per slot and tet word
  int sd[3][2][32] = {};        // group shift d: tet bit q takes triangle bit q - d
  uint32_t sm[3][2][32] = {};   // tet bits of the group
  int nug[2] = {};              // uncertainty groups per tet word (slots merged)
  int ud[2][32] = {};
  uint32_t um[2][32] = {};
  uint32_t axis[4][2] = {};     // tets whose last corner raises axis j
  uint8_t first[64] = {};       // first tet (a, b, c) of triangle (a, b), and their count
  uint8_t count[64] = {};
};

template <int D>
constexpr SliceTab<D> build_slice_tab() {
  SliceTab<D> S{};
  const auto C2 = build_chains<D, 2>();
  const auto C3 = build_chains<D, 3>();
  auto tri = [&](int a, int b) {
    for (int t = 0; t < C2.n; ++t)
      if (C2.s[t][1] == a && C2.s[t][2] == b) return t;
    return -1;
  };
  for (int k = 0; k < C3.n; ++k) {
    const int a = C3.s[k][1], b = C3.s[k][2], c = C3.s[k][3];
    const int w = k >> 5, q = push_bit(k, C3.n);
    const int t3[3] = {tri(a, b), tri(a, c), tri(b, c)};
    for (int s = 0; s < 3; ++s) {
      const int p = 32 * (t3[s] >> 5) + push_bit(t3[s], C2.n);
      const int d = q - p;
      int g = 0;
      while (g < S.nsg[s][w] && S.sd[s][w][g] != d) ++g;
      if (g == S.nsg[s][w]) S.sd[s][w][S.nsg[s][w]++] = d;
      S.sm[s][w][g] |= 1u << q;
      int u = 0;
      while (u < S.nug[w] && S.ud[w][u] != d) ++u;
      if (u == S.nug[w]) S.ud[w][S.nug[w]++] = d;
      S.um[w][u] |= 1u << q;
    }
    for (int j = 0; j < D; ++j)
      if ((c >> j) & 1) S.axis[j][w] |= 1u << q;
    const int ab = t3[0];
    if (S.count[ab] == 0) S.first[ab] = (uint8_t)k;
    S.count[ab]++;
  }
  return S;
}
template <int D>
constexpr SliceTab<D> SLICE = build_slice_tab<D>();

// Bits [q - d] of the 64-bit word, aligned to bit q (a compile-time funnel shift).
template <int d>
__device__ __forceinline__ uint32_t shifted64(uint64_t w) {
  if constexpr (d >= 0)
    return (uint32_t)(w << d);
  else
    return (uint32_t)(w >> -d);
}

// Items pushed one bit at a time (push_bit layout) back to bit i = item i.
template <int N>
__device__ __forceinline__ uint64_t unpush(const uint32_t* w) {
  if constexpr (N <= 32) {
    return (uint64_t)(__builtin_bitreverse32(w[0]) >> (32 - N));
  } else {
    return (uint64_t)__builtin_bitreverse32(w[0]) |
           ((uint64_t)(__builtin_bitreverse32(w[1]) >> (64 - N)) << 32);
  }
}

// ---------------------------------------------------------------------------
// Per-vertex classification, two stages.
//
// Stage 1 (classify_fast, fully unrolled, branch-light): for every simplex type the
// vertex owns, decide from exact level-0 signs whether it is rejected, inside by
// SoS (needs the solver gate for critical points), or degenerate (some level-0
// value is exactly zero, or the magnitudes exceed the fast filter: needs the
// verbatim multi-level sign tests).  Stage 2 spreads those per-simplex jobs over
// the workgroup's lanes (run_items) so every lane does useful work.
//
// Bits: acc = accepted, sol = SoS inside, solver pending, deg = verbatim pending.
// A triple product whose filtered FP32 value is within its error bound (exact zeros
// included) sends the tet to the verbatim path, which decides it exactly.

struct Masks {
  uint64_t acc, sol, deg;
  uint64_t tdeg = 0;  // degenerate simplices decidable from the stage-1 LDS tile (M = 3)
};

// LDS tile of quantized 3-component values.  A workgroup owns the 256 consecutive
// vertices [v0, v0 + 256); the corners of all of them lie in 2^(D-1) runs of 257
// consecutive vertices (one per combination of the non-x axes), each loaded
// coalesced and quantized once.  Values are exact integers held in float (the
// fast path requires |q| < 2^20).
template <int D>
struct Tile3 {
  static constexpr int NSEG = 1 << (D - 1);
  static constexpr int SEG = BLOCK + 1;
  static constexpr int FLOATS = NSEG * SEG * 3;
};

// Loads of one LDS tile: element e is vertex e % SEG of run e / SEG.  Every load of
// the thread's elements is issued before any value is used (a loop that stores each
// element before loading the next waits for memory once per element).  Element e of
// unrolled step `it` (e = thread + it * BLOCK) lies in run it or it - 1, so the run
// offset is one of two compile-time kernel-argument entries.
template <class T, int D, int NCOMP, class F>
__device__ __forceinline__ void stage_tile(const KParams& P, int64_t v0, F&& store) {
  constexpr int NSEG = Tile3<D>::NSEG, SEG = Tile3<D>::SEG, NE = NSEG * SEG;
  constexpr int IT = (NE + BLOCK - 1) / BLOCK;
  if (!P.packed) {
    // strided views: one element at a time (the index decomposition dominates)
    for (int e = threadIdx.x; e < NE; e += BLOCK) {
      const int seg = e / SEG;
      int64_t u = v0 + P.loff[seg << 1] + (e - seg * SEG);
      u = u < P.nvert ? u : P.nvert - 1;  // such corners lie outside the grid and are masked
      int64_t l[4];
      decompose<D>(P, u, l);
      const VtxRef r = make_ref<D>(P, u, l);
      T v[NCOMP];
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) v[k] = load_val<T>(P, k, r, 0);
      store(e, v);
    }
    return;
  }
  T val[IT][NCOMP];
  sfor<0, IT>([&](auto ii) {
    constexpr int it = decltype(ii)::value;
    constexpr int s_hi = it < NSEG ? it : NSEG - 1, s_lo = it > 0 ? it - 1 : 0;
    const int e = (int)threadIdx.x + it * BLOCK;
    if (e < NE) {
      const bool lo = (int)threadIdx.x < it;  // e < it * SEG
      const int seg = lo ? s_lo : s_hi;
      int64_t u = v0 + (lo ? P.loff[s_lo << 1] : P.loff[s_hi << 1]) + (e - seg * SEG);
      u = u < P.nvert ? u : P.nvert - 1;  // such corners lie outside the grid and are masked
      VtxRef r;
      r.base[0] = r.base[1] = r.base[2] = u;
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) val[it][k] = load_val<T>(P, k, r, 0);
    }
  });
  sfor<0, IT>([&](auto ii) {
    constexpr int it = decltype(ii)::value;
    const int e = (int)threadIdx.x + it * BLOCK;
    if (e < NE) store(e, val[it]);
  });
}

template <class T, int D>
__device__ __forceinline__ void stage_quantized3(const KParams& P, int64_t v0, float* tile) {
  stage_tile<T, D, 3>(P, v0, [&](int e, const T* v) {
#pragma unroll
    for (int k = 0; k < 3; ++k) tile[e * 3 + k] = (float)qt(v[k], P.Q);
  });
}

// The same tile for the det2 predicates (CriticalPointPredicate<2,T>, and
// FiberPredicate<T> on values minus thresholds): two quantized values per vertex as
// exact integer-valued doubles, each corner value loaded and quantized once per
// workgroup instead of once per vertex whose hypercube uses it.
template <int KIND, class T, int D>
__device__ __forceinline__ void stage_quantized2(const KParams& P, int64_t v0, double* tile) {
  stage_tile<T, D, 2>(P, v0, [&](int e, const T* v) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      T w = v[k];
      if constexpr (KIND == KIND_FIBER) w = w - (T)P.thr[k];
      tile[e * 2 + k] = qt(w, P.Q);
    }
  });
}

// Critical points of a 3-component field on 3-simplices (CriticalPointPredicate<3,T>,
// predicate.hpp:66-99).
template <class T, int D>
__device__ __forceinline__ Masks classify_cp3(const KParams& P, const Vtx& x, const float* tile) {
  constexpr int NC = 1 << D, FULL = NC - 1;
  const uint32_t inb = corner_mask<D>(P, x.l);
  // Corners outside the grid read some other vertex's value from the tile; a tet
  // whose last corner c is inside has all its corners inside (they are subsets of
  // c), so only the tet masks need the in-grid gate, and their values only raise mx
  // (a larger filter bound is still a bound).
  float X[NC][3];
  float mx = 0.f;  // the largest |q|, exact (integers below 2^24 are exact in float)
#pragma unroll
  for (int s = 0; s < NC; ++s) {
    const float* src = tile + ((s >> 1) * Tile3<D>::SEG + threadIdx.x + (s & 1)) * 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      X[s][k] = src[k];
      mx = fmaxf(mx, fabsf(X[s][k]));
    }
  }
  if (!(mx < 0x1p20f)) {
    // magnitudes beyond the fast filter: every tet takes the verbatim path
    uint64_t deg = 0;
    sfor<0, CH<D, 3>.n>([&](auto tt) {
      constexpr int t = decltype(tt)::value;
      constexpr int c = CH<D, 3>.s[t][3];
      if ((inb >> c) & 1) deg |= 1ull << t;
    });
    return Masks{0, 0, deg};
  }
  const float bnd = mx * mx * mx * FTK2H_TP32_BOUND;  // mx^3 < 2^60; rounding well inside the 2x margin
  // Each triple product f gives two bits, shifted into 32-bit words in the order
  // computed (one v_alignbit each): its sign bit (the exact sign when f is certain)
  // and "certain" = the sign bit of bnd - |f|, set iff |f| > bnd (a difference of
  // two floats is zero only when they are equal and keeps its sign when flushed).
  // tp(0,a,b) = (Xa x Xb) . X0 and tp(a,b,c) = (Xa x Xb) . Xc share the cross
  // product of (a, b); for b = FULL (no tet has it as its middle corner)
  // tp(0,a,FULL) = (X_FULL x X0) . Xa.
  constexpr int NTRI = CH<D, 2>.n, NTET = CH<D, 3>.n;
  uint32_t tn[2] = {0, 0}, tc[2] = {0, 0};  // triangles (0, a, b), by type
  uint32_t fn[2] = {0, 0}, fc[2] = {0, 0};  // tets' opposite faces (a, b, c), by type
  float ef[3];
  cross3f(X[FULL], X[0], ef);
  sfor<0, NTRI>([&](auto tt) {
    constexpr int t = decltype(tt)::value;
    constexpr int a = CH<D, 2>.s[t][1], b = CH<D, 2>.s[t][2];
    float e[3];
    float f;
    if constexpr (b == FULL) {
      f = dot3f(ef, X[a]);
    } else {
      cross3f(X[a], X[b], e);
      f = dot3f(e, X[0]);
    }
    tn[t >> 5] = __builtin_amdgcn_alignbit(tn[t >> 5], __float_as_uint(f), 31);
    tc[t >> 5] = __builtin_amdgcn_alignbit(tc[t >> 5], __float_as_uint(bnd - fabsf(f)), 31);
    if constexpr (b != FULL) {
      sfor<SLICE<D>.first[t], SLICE<D>.first[t] + SLICE<D>.count[t]>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int c = CH<D, 3>.s[k][3];
        const float g = dot3f(e, X[c]);
        fn[k >> 5] = __builtin_amdgcn_alignbit(fn[k >> 5], __float_as_uint(g), 31);
        fc[k >> 5] = __builtin_amdgcn_alignbit(fc[k >> 5], __float_as_uint(bnd - fabsf(g)), 31);
      });
    }
    __builtin_amdgcn_sched_barrier(0);  // keep live ranges short
  });
  // Gather each tet's three triangle bits to its own position (see SliceTab).
  const uint64_t sn = ((uint64_t)tn[1] << 32) | tn[0];
  const uint64_t su = ~(((uint64_t)tc[1] << 32) | tc[0]);  // uncertain triangles
  constexpr int NW = (NTET + 31) / 32;
  uint32_t nw[3][2] = {}, uw[2] = {};
  sfor<0, NW>([&](auto ww) {
    constexpr int w = decltype(ww)::value;
    sfor<0, 3>([&](auto ss) {
      constexpr int s = decltype(ss)::value;
      sfor<0, SLICE<D>.nsg[s][w]>([&](auto gg) {
        constexpr int g = decltype(gg)::value;
        constexpr uint32_t m = SLICE<D>.sm[s][w][g];
        nw[s][w] |= shifted64<SLICE<D>.sd[s][w][g]>(sn) & m;
      });
    });
    sfor<0, SLICE<D>.nug[w]>([&](auto gg) {
      constexpr int g = decltype(gg)::value;
      constexpr uint32_t m = SLICE<D>.um[w][g];
      uw[w] |= shifted64<SLICE<D>.ud[w][g]>(su) & m;
    });
  });
  // d3 = det3(v0,v1,v2) = tp(0ab); d2 = det3(v0,v3,v1) = -tp(0ac);
  // d1 = det3(v0,v2,v3) = tp(0bc); d0 = det3(v1,v3,v2) = -tp(abc).
  // Inside by level 0 iff all four are equal; any uncertain/zero value -> verbatim.
  // With all four certain, each comparison is an XOR of two sign bits:
  // d3 == d2 <=> n012 != n013;  d1 == d3 <=> n023 == n012;  d0 == d3 <=> nabc != n012.
  uint32_t solw[2] = {0, 0}, degw[2] = {0, 0};
  sfor<0, NW>([&](auto ww) {
    constexpr int w = decltype(ww)::value;
    uint32_t invalid = 0;
    sfor<0, D>([&](auto jj) {
      constexpr int j = decltype(jj)::value;
      constexpr uint32_t ax = SLICE<D>.axis[j][w];
      invalid |= ((inb >> (1 << j)) & 1) ? 0u : ax;
    });
    const uint32_t n012 = nw[0][w], n013 = nw[1][w], n023 = nw[2][w], nabc = fn[w];
    const uint32_t in = (n012 ^ n013) & ~(n023 ^ n012) & (nabc ^ n012);
    const uint32_t cert = ~uw[w] & fc[w];
    solw[w] = ~invalid & cert & in;
    degw[w] = ~invalid & ~cert;
  });
  // the verbatim levels run from the tile
  return Masks{0, unpush<NTET>(solw), 0, unpush<NTET>(degw)};
}

// det2-based predicates on 2-simplices: CriticalPointPredicate<2,T> and
// FiberPredicate<T> (values minus thresholds, predicate.hpp:352).
template <int KIND, class T, int D>
__device__ __forceinline__ T val2(const KParams& P, int k, const Vtx& x, int s) {
  const T v = load_val<T>(P, k, x.r, s);
  if constexpr (KIND == KIND_FIBER) return v - (T)P.thr[k];
  return v;
}

// Whether stage 1 of a det2 predicate reads its corners from the LDS tile
// (stage_quantized2): on 2-D spacetime meshes (4 segments, 16 KB); on 3-D ones
// direct loads measured faster than a tile of 8 segments, as doubles (33 KB) or as
// floats exact below 2^24 (16 KB; fiber 256^3 x 4 stage 1 2.08 -> 2.23 ms).
template <int D, int M>
constexpr bool TILE2 = M == 2 && D == 3;

__device__ __forceinline__ double cross2(const double* A, const double* B) {
  // exact for |q| < 2^26: products < 2^52
  return fma(A[0], B[1], -(A[1] * B[0]));
}

template <int KIND, class T, int D>
__device__ __forceinline__ Masks classify_tri(const KParams& P, const Vtx& x, const double* tile) {
  uint64_t acc = 0, sol = 0, deg = 0;
  constexpr int NC = 1 << D;
  const uint32_t inb = corner_mask<D>(P, x.l);
  double X[NC][2];
  double mx = 0;
#pragma unroll
  for (int s = 0; s < NC; ++s) {
    const bool in = (inb >> s) & 1;
    const double* src = tile + ((s >> 1) * Tile3<D>::SEG + threadIdx.x + (s & 1)) * 2;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      double q;
      if constexpr (TILE2<D, 2>) {
        q = in ? src[k] : 0.0;  // qt(value - threshold) of corner s
      } else {
        const T v = val2<KIND, T, D>(P, k, x, in ? s : 0);
        q = in ? qt(v, P.Q) : 0.0;
      }
      mx = fmax(mx, fabs(q));
      X[s][k] = q;
    }
  }
  const bool slow = !(mx < 0x1p26);
  sfor<0, CH<D, 2>.n>([&](auto tt) {
    constexpr int t = decltype(tt)::value;
    constexpr int a = CH<D, 2>.s[t][1], b = CH<D, 2>.s[t][2];
    const uint64_t valid = (inb >> b) & 1;
    const double c01 = cross2(X[0], X[a]);
    const double c12 = cross2(X[a], X[b]);
    const double c20 = cross2(X[b], X[0]);
    const bool s0 = c01 < 0;   // det2 tie: i0 < i1 -> +1
    const bool s1 = c12 < 0;   // tie: i1 < i2 -> +1
    const bool s2 = !(c20 > 0);  // tie: i2 > i0 -> -1
    const uint64_t in = (uint64_t)(s0 == s1 && s1 == s2);
    deg |= (valid & (uint64_t)slow) << t;
    const uint64_t hit = valid & (uint64_t)!slow & in;
    if constexpr (KIND == KIND_CP)
      sol |= hit << t;
    else
      acc |= hit << t;  // fiber: solver failure falls back to 1/3
  });
  return Masks{acc, sol, deg};
}

// Edge predicates: ContourPredicate<T> (sign change of value - threshold) and
// CriticalPointPredicate<1,T> (sign change + solver gate).
// The sign class of one corner value for the edge predicates, staged once per
// workgroup (stage_sign_classes): 1 / -1 when the quantized value is at least one
// quantum from zero, 0 for a tie, whose sos_sign comes from the corner's global id.
template <int KIND, class T>
__device__ __forceinline__ int8_t sign_class(const KParams& P, T v) {
  if constexpr (KIND == KIND_CONTOUR) v = v - (T)P.thr[0];
  const double r = qd(v, P.Q);  // as sos_sign
  return r >= 1.0 ? (int8_t)1 : (r <= -1.0 ? (int8_t)-1 : (int8_t)0);
}
template <int KIND, class T, int D>
__device__ __forceinline__ void stage_sign_classes(const KParams& P, int64_t v0, int8_t* cls) {
  stage_tile<T, D, 1>(P, v0, [&](int e, const T* v) { cls[e] = sign_class<KIND, T>(P, v[0]); });
}

template <int KIND, class T, int D>
__device__ __forceinline__ Masks classify_edge(const KParams& P, const Vtx& x, const int8_t* cls) {
  uint64_t acc = 0, sol = 0, deg = 0;
  constexpr int NC = 1 << D;
  const uint32_t inb = corner_mask<D>(P, x.l);
  int sg[NC];
#pragma unroll
  for (int s = 0; s < NC; ++s) {
    const bool in = (inb >> s) & 1;
    const int c = cls[(s >> 1) * Tile3<D>::SEG + threadIdx.x + (s & 1)];
    const int tie = ((x.r.gid + P.goff[s]) % 2 == 0) ? 1 : -1;
    sg[s] = in ? (c != 0 ? c : tie) : 1;
  }
#pragma unroll
  for (int s = 1; s < NC; ++s) {
    // Chains<D,1> type of edge (0, s) is s - 1
    const uint64_t hit = ((inb >> s) & 1) & (uint64_t)(sg[s] != sg[0]);
    if constexpr (KIND == KIND_CP)
      sol |= hit << (s - 1);
    else
      acc |= hit << (s - 1);
  }
  return Masks{acc, sol, deg};
}

template <int KIND, int D, int M, class T>
__device__ __forceinline__ Masks classify_fast(const KParams& P, const Vtx& x, const float* tile) {
  if constexpr (M == 3)
    return classify_cp3<T, D>(P, x, tile);
  else if constexpr (M == 2)
    return classify_tri<KIND, T, D>(P, x, reinterpret_cast<const double*>(tile));
  else
    return classify_edge<KIND, T, D>(P, x, reinterpret_cast<const int8_t*>(tile));
}

// Stage 2 helpers for one pending simplex (type t owned by the vertex of r).
// Field values of its vertices (fiber: minus the thresholds).
template <int KIND, int D, int M, class T>
__device__ __forceinline__ void item_values(const KParams& P, const CornerTab& tab, const VtxRef& r, int t,
                                            T (*vals)[M]) {
  const uint32_t pc = tab.pc[t];
  if (P.packed) {
    // every component x-fastest with the local extents: one element index per corner
    int64_t e[M + 1];
#pragma unroll
    for (int i = 0; i <= M; ++i) {
      const int off = i == 0 ? 0 : (int)((pc >> (4 * (i - 1))) & 0xF);
      e[i] = r.base[0] + tab.coff[0][off];
    }
#pragma unroll
    for (int i = 0; i <= M; ++i)
#pragma unroll
      for (int k = 0; k < M; ++k) {
        T v = gptr(static_cast<const T*>(P.comp[k]))[e[i]];
        if constexpr (KIND == KIND_FIBER) v = v - (T)P.thr[k];
        vals[i][k] = v;
      }
    return;
  }
#pragma unroll
  for (int i = 0; i <= M; ++i) {
    const int off = i == 0 ? 0 : (int)((pc >> (4 * (i - 1))) & 0xF);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      T v = load_val_t<T>(P, tab, k, r, off);
      if constexpr (KIND == KIND_FIBER) v = v - (T)P.thr[k];
      vals[i][k] = v;
    }
  }
}

// The verbatim SoS test (origin_inside<M>, sos.hpp:202-254) of a degenerate simplex.
template <int KIND, int D, int M, class T>
__device__ __forceinline__ bool item_inside(const KParams& P, const CornerTab& tab, const VtxRef& r, int t) {
  static_assert(M >= 2, "edges never take the verbatim path");
  T vals[M + 1][M];
  item_values<KIND, D, M, T>(P, tab, r, t, vals);
  double Xd[M + 1][M];
  double mx = 0;
#pragma unroll
  for (int i = 0; i <= M; ++i)
#pragma unroll
    for (int k = 0; k < M; ++k) {
      Xd[i][k] = qt(vals[i][k], P.Q);
      mx = fmax(mx, fabs(Xd[i][k]));
    }
  if constexpr (M == 3) {
    if (mx < 0x1p20) return inside3_exact_d(Xd, mx);
    int64_t X[M + 1][M];
    for (int i = 0; i <= M; ++i)
      for (int k = 0; k < M; ++k) X[i][k] = q64(vals[i][k], P.Q);
    return inside3_exact(X);
  } else {
    int64_t X[M + 1][M];
#pragma unroll
    for (int i = 0; i <= M; ++i)
#pragma unroll
      for (int k = 0; k < M; ++k) X[i][k] = q64(vals[i][k], P.Q);
    return inside2_exact(X);
  }
}

// The solver gate of CriticalPointPredicate (predicate.hpp:74): a failed solve
// rejects the simplex.
template <int KIND, int D, int M, class T>
__device__ __forceinline__ bool item_solves(const KParams& P, const CornerTab& tab, const VtxRef& r, int t) {
  if constexpr (M == 1) {
    const uint32_t pc = tab.pc[t];
    T v0 = load_val_t<T>(P, tab, 0, r, 0);
    T v1 = load_val_t<T>(P, tab, 0, r, (int)(pc & 0xF));
    T lam[2];
    return solve_1<T>(v0, v1, lam);
  } else {
    T vals[M + 1][M];
    item_values<KIND, D, M, T>(P, tab, r, t, vals);
    T lam[M + 1];
    return solve_k<M, T>(vals, lam);
  }
}

// ---------------------------------------------------------------------------
// Values of one node's vertices (type t owned by the vertex of r): vals[i][k] =
// component k at vertex i (fiber and contour: minus the thresholds).
template <int KIND, int D, int M, class T>
__device__ __forceinline__ void node_values(const KParams& P, const CornerTab& tab, const VtxRef& r, int t,
                                            T (*vals)[M]) {
  const uint32_t pc = tab.pc[t];
  if (P.packed) {
    // every component x-fastest with the local extents: one element index per corner
    int64_t e[M + 1];
#pragma unroll
    for (int i = 0; i <= M; ++i) {
      const int off = i == 0 ? 0 : (int)((pc >> (4 * (i - 1))) & 0xF);
      e[i] = r.base[0] + tab.coff[0][off];
    }
#pragma unroll
    for (int i = 0; i <= M; ++i)
#pragma unroll
      for (int k = 0; k < M; ++k) {
        T v = gptr(static_cast<const T*>(P.comp[k]))[e[i]];
        if constexpr (KIND == KIND_FIBER || KIND == KIND_CONTOUR) v = v - (T)P.thr[k];
        vals[i][k] = v;
      }
    return;
  }
#pragma unroll
  for (int i = 0; i <= M; ++i) {
    const int off = i == 0 ? 0 : (int)((pc >> (4 * (i - 1))) & 0xF);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      T v = load_val_t<T>(P, tab, k, r, off);
      if constexpr (KIND == KIND_FIBER || KIND == KIND_CONTOUR) v = v - (T)P.thr[k];
      vals[i][k] = v;
    }
  }
}

// lambda of one node from its vertex values, written as T[M+1]; returns scalar.
template <int KIND, int D, int M, class T>
__device__ __forceinline__ void node_solve(const KParams& P, const CornerTab& tab, const VtxRef& r, int t,
                                           const T (*vals)[M], T* lam, float* scalar) {
  *scalar = 0.f;
  if constexpr (M == 1) {
    if constexpr (KIND == KIND_CONTOUR) {
      if (!solve_1<T>(vals[0][0], vals[1][0], lam)) lam[0] = lam[1] = (T)0.5;
      *scalar = (float)(T)P.thr[0];
    } else {
      solve_1<T>(vals[0][0], vals[1][0], lam);
    }
  } else {
    const bool ok = solve_k<M, T>(vals, lam);
    if constexpr (KIND == KIND_FIBER) {
      if (!ok)
#pragma unroll
        for (int i = 0; i <= M; ++i) lam[i] = (T)(1.0 / 3.0);
    }
  }
  if constexpr (KIND == KIND_CP) {
    if (P.scal) {  // predicate.hpp:81-92
      const uint32_t pc = tab.pc[t];
      T sv = T(0);
#pragma unroll
      for (int i = 0; i <= M; ++i) {
        const int off = i == 0 ? 0 : (int)((pc >> (4 * (i - 1))) & 0xF);
        sv = sv + lam[i] * gptr(static_cast<const T*>(P.scal))[r.sbase + tab.soff[off]];
      }
      *scalar = (float)sv;
    }
  }
}

// ---------------------------------------------------------------------------
// block-wide exclusive scan of one uint32 per thread (BLOCK threads, wave64)
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* total) {
  __shared__ uint32_t wsum[BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) {
    const uint32_t s = wsum[i];
    if (i < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + incl - x;
}

// ---------------------------------------------------------------------------
// Manifold cells of the hypercube based at l (all l_j <= dims_j - 2).
//   Mode 0: count only.  Mode 1: write cells (canonical indices, each cell sorted).
template <int D, int M>
struct OwnerSet {
  static constexpr uint32_t bits() {
    uint32_t b = 0;
    for (int p = 0; p < CellFaceTable<D, M>::NP; ++p)
      for (int f = 0; f < CellFaceTable<D, M>::NF; ++f) b |= 1u << CF<D, M>.owner[p][f];
    return b;
  }
};

template <int KIND, int D, int M>
struct CellKind {
  // 1: star edges (k = 1); 2: star-fan triangles (k = 2); 3: marching pentatope
  // (contour, d = 4); 4: marching tetrahedron (contour, d = 3); 0: none.
  static constexpr int value = (M == 1 && (D == 4 || D == 3))
                                   ? (KIND == KIND_CONTOUR ? (D == 4 ? 3 : 4) : 0)
                                   : ((D - M) == 1 ? 1 : ((D - M) == 2 ? 2 : 0));
  static constexpr int verts = value == 1 ? 2 : (value == 2 ? 3 : (value == 3 ? 4 : (value == 4 ? 3 : 0)));
  static constexpr int dim = verts - 1;
};

template <int N>
__device__ __forceinline__ void sort_small(uint32_t* c) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j)
      if (c[j] < c[i]) {
        const uint32_t t = c[i];
        c[i] = c[j];
        c[j] = t;
      }
}

template <int KIND, int D, int M, class Ptr = uint32_t*>
__device__ __forceinline__ uint32_t cells_of(const KParams& P, const Vtx& x, int mode, Ptr dst) {
  constexpr int CK = CellKind<KIND, D, M>::value;
  if constexpr (CK == 0) {
    return 0;
  } else {
    constexpr int NV = CellKind<KIND, D, M>::verts;
    constexpr int NP = CellFaceTable<D, M>::NP;
    constexpr int NF = CellFaceTable<D, M>::NF;
    constexpr uint32_t OWN = OwnerSet<D, M>::bits();
    uint64_t mk[1 << D];
    uint32_t no[1 << D];
#pragma unroll
    for (int o = 0; o < (1 << D); ++o) {
      mk[o] = 0;
      no[o] = 0;
      if ((OWN >> o) & 1) {
        const int64_t u = x.v + P.loff[o];
        mk[o] = gptr(P.mask)[u];
        if (mode) no[o] = gptr(P.node_off)[u];
      }
    }
    auto idx = [&](int o, int t) -> uint32_t {
      return no[o] + (uint32_t)__popcll(mk[o] & ((1ull << t) - 1ull));
    };
    uint32_t count = 0;
    auto put = [&](uint32_t* c) {
      if (mode) {
        sort_small<NV>(c);
#pragma unroll
        for (int i = 0; i < NV; ++i) dst[count * NV + i] = c[i];
      }
      count++;
    };
    if constexpr (CK == 1 || CK == 2) {
      sfor<0, NP>([&](auto pp) {
        constexpr int p = decltype(pp)::value;
        // active faces in encode_simplex_id order
        uint32_t act = 0;  // bit k: k-th face in order is active
        sfor<0, NF>([&](auto kk) {
          constexpr int k = decltype(kk)::value;
          constexpr int f = CF<D, M>.order[p][k];
          constexpr int o = CF<D, M>.owner[p][f], t = CF<D, M>.type[p][f];
          if ((mk[o] >> t) & 1) act |= 1u << k;
        });
        if (act == 0) return;
        const int k0 = __builtin_ctz(act);
        if constexpr (CK == 1) {
          if (mode == 0) {
            count += __builtin_popcount(act) - 1;
          } else {
            uint32_t hs = 0;
            sfor<0, NF>([&](auto kk) {
              constexpr int k = decltype(kk)::value;
              constexpr int f = CF<D, M>.order[p][k];
              constexpr int o = CF<D, M>.owner[p][f], ty = CF<D, M>.type[p][f];
              if (k == k0) hs = idx(o, ty);
            });
            sfor<0, NF>([&](auto kk) {
              constexpr int k = decltype(kk)::value;
              constexpr int f = CF<D, M>.order[p][k];
              if (((act >> k) & 1) && k != k0) {
                constexpr int o = CF<D, M>.owner[p][f], ty = CF<D, M>.type[p][f];
                uint32_t c[2] = {hs, idx(o, ty)};
                put(c);
              }
            });
          }
        } else {
          // star-fan (engine.hpp:1388-1415): for each cell facet (drop vertex i)
          uint32_t hs = 0;
          if (mode) {
            sfor<0, NF>([&](auto kk) {
              constexpr int k = decltype(kk)::value;
              constexpr int f = CF<D, M>.order[p][k];
              constexpr int o = CF<D, M>.owner[p][f], ty = CF<D, M>.type[p][f];
              if (k == k0) hs = idx(o, ty);
            });
          }
          sfor<0, D + 1>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            // faces (in order) that avoid cell vertex i
            uint32_t onf = 0;
            sfor<0, NF>([&](auto kk) {
              constexpr int k = decltype(kk)::value;
              constexpr int f = CF<D, M>.order[p][k];
              constexpr int vm = CF<D, M>.vmask[f];
              if (!((vm >> i) & 1)) onf |= 1u << k;
            });
            const uint32_t sel = act & onf;
            if (__builtin_popcount(sel) < 2) return;
            const int kt = __builtin_ctz(sel);
            // h_S != h_T: the star centre must not be the facet's first node
            if (kt == k0) return;
            // every other node n of the facet with n != h_S
            const uint32_t rest = sel & ~(1u << kt) & ~(1u << k0);
            if (mode == 0) {
              count += __builtin_popcount(rest);
            } else {
              uint32_t ht = 0;
              sfor<0, NF>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int f = CF<D, M>.order[p][k];
                constexpr int o = CF<D, M>.owner[p][f], ty = CF<D, M>.type[p][f];
                if (k == kt) ht = idx(o, ty);
              });
              sfor<0, NF>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int f = CF<D, M>.order[p][k];
                if ((rest >> k) & 1) {
                  constexpr int o = CF<D, M>.owner[p][f], ty = CF<D, M>.type[p][f];
                  uint32_t c[3] = {hs, ht, idx(o, ty)};
                  put(c);
                }
              });
            }
          });
        }
      });
    } else {
      // marching pentatope / tetrahedron on vertex signs relative to vertex 0 of
      // the cell: edge (0, s) of hypercube v is active iff the signs differ.
      constexpr int NPC = CellFaceTable<D, 1>::NP;
      sfor<0, NPC>([&](auto pp) {
        constexpr int p = decltype(pp)::value;
        int rel[D + 1];
        int ce[D + 1];  // cell vertex offsets (compile-time constants)
        sfor<0, D + 1>([&](auto qq) {
          constexpr int q = decltype(qq)::value;
          constexpr int cq = CF<D, 1>.cell[p][q];
          ce[q] = cq;
          rel[q] = q == 0 ? 0 : (int)((mk[0] >> (cq > 0 ? cq - 1 : 0)) & 1);
        });
        int g1 = 0;
#pragma unroll
        for (int i = 0; i <= D; ++i) g1 += rel[i];
        const int g0 = D + 1 - g1;
        // node of cell edge (i, j), i < j: owner c_i, type c_j ^ c_i - 1
        auto eidx = [&](int i, int j) -> uint32_t {
          int ci = 0, cj = 0;
#pragma unroll
          for (int q = 0; q <= D; ++q) {
            if (q == i) ci = ce[q];
            if (q == j) cj = ce[q];
          }
          const int lo = i < j ? ci : cj, hi = i < j ? cj : ci;
          uint32_t r = 0;
#pragma unroll
          for (int o = 0; o < (1 << D); ++o)
            if (o == lo) r = idx(o, (lo ^ hi) - 1);
          return r;
        };
        // groups in position order; grp[0] holds vertex 0's side.  Indices are
        // written through compile-time selects so the arrays stay in registers.
        int grp[2][D + 1];
        int gn0 = 0, gn1 = 0;
#pragma unroll
        for (int i = 0; i <= D; ++i) {
#pragma unroll
          for (int q = 0; q <= D; ++q) {
            if (rel[i] == 0 && q == gn0) grp[0][q] = i;
            if (rel[i] == 1 && q == gn1) grp[1][q] = i;
          }
          if (rel[i] == 0)
            gn0++;
          else
            gn1++;
        }
        if constexpr (CK == 3) {
          if (g0 == 1 || g1 == 1) {
            const int sg = (g0 == 1) ? 0 : 1;
            if (mode == 0) {
              count += 1;
            } else {
              uint32_t c[4];
#pragma unroll
              for (int q = 0; q < 4; ++q)
                c[q] = sg == 0 ? eidx(grp[0][0], grp[1][q]) : eidx(grp[1][0], grp[0][q]);
              put(c);
            }
          } else if (g0 == 2 || g1 == 2) {
            if (mode == 0) {
              count += 3;
            } else {
              const int tw = (g0 == 2) ? 0 : 1;
              uint32_t T0[3], T1[3];
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                T0[q] = tw == 0 ? eidx(grp[0][0], grp[1][q]) : eidx(grp[1][0], grp[0][q]);
                T1[q] = tw == 0 ? eidx(grp[0][1], grp[1][q]) : eidx(grp[1][1], grp[0][q]);
              }
              uint32_t c0[4] = {T0[0], T0[1], T0[2], T1[2]};
              uint32_t c1[4] = {T0[0], T0[1], T1[1], T1[2]};
              uint32_t c2[4] = {T0[0], T1[0], T1[1], T1[2]};
              put(c0);
              put(c1);
              put(c2);
            }
          }
        } else {  // CK == 4: marching tetrahedron (d = 3)
          if (g0 == 1 || g1 == 1) {
            if (mode == 0) {
              count += 1;
            } else {
              const int sg = (g0 == 1) ? 0 : 1;
              uint32_t c[3];
#pragma unroll
              for (int q = 0; q < 3; ++q)
                c[q] = sg == 0 ? eidx(grp[0][0], grp[1][q]) : eidx(grp[1][0], grp[0][q]);
              put(c);
            }
          } else if (g0 == 2 && g1 == 2) {
            if (mode == 0) {
              count += 2;
            } else {
              uint32_t n[4];
              int k = 0;
#pragma unroll
              for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) n[k++] = eidx(grp[0][a], grp[1][b]);
              sort_small<4>(n);
              uint32_t c0[3] = {n[0], n[1], n[2]};
              uint32_t c1[3] = {n[0], n[2], n[3]};
              put(c0);
              put(c1);
            }
          }
        }
      });
    }
    return count;
  }
}

// ---------------------------------------------------------------------------
// Kernels
// Tile functions: one workgroup processes the 256 vertices [blk*256, blk*256 + 256)
// of the mesh described by P.  The single-mesh kernels run one tile per workgroup;
// the batched kernels (batch.hip) loop over tiles of many meshes.
// Position of the n-th set bit of m (n < popcount(m)), by halving.
__device__ __forceinline__ int nth_set_bit(uint64_t m, uint32_t n) {
  int pos = 0;
  uint32_t x = (uint32_t)m;
  uint32_t c = __popc(x);
  if (n >= c) {
    n -= c;
    x = (uint32_t)(m >> 32);
    pos = 32;
  }
#pragma unroll
  for (int h = 16; h >= 1; h >>= 1) {
    c = __popc(x & ((1u << h) - 1u));
    if (n >= c) {
      n -= c;
      x >>= h;
      pos += h;
    }
  }
  return pos;
}

// The items of a tile: thread l owns the set bits of its mask w_l.  Items are
// numbered in (thread, bit) order; item j belongs to the thread l with
// offs[l] <= j < offs[l+1] (offs = exclusive scan of the popcounts) and is that
// thread's (j - offs[l])-th set bit.  Consecutive threads take consecutive items,
// so every lane does useful work however the bits are spread, without building a
// queue.
// With Q > 0 a tile's items are also listed directly (owner thread << 6 | type) when
// there are at most Q of them (stage 2: 3.11 -> 2.96 ms); otherwise item j is found
// by a binary search (stage 1's degenerate tets: the list's writes cost more than
// the searches they save).
template <int Q>
struct ItemLds {
  uint32_t offs[BLOCK + 1];
  unsigned long long masks[BLOCK];
  uint16_t q[Q > 0 ? Q : 1];
};

template <int Q>
__device__ __forceinline__ void item_at(const ItemLds<Q>& L, uint32_t tot, uint32_t j, int& owner, int& t) {
  if (Q > 0 && tot <= (uint32_t)Q) {
    const uint32_t e = L.q[j];
    owner = (int)(e >> 6);
    t = (int)(e & 63);
    return;
  }
  int lo = 0, hi = BLOCK;  // offs[lo] <= j < offs[hi]
#pragma unroll
  for (int s = 0; s < 8; ++s) {  // BLOCK = 256
    const int mid = (lo + hi) >> 1;
    if (L.offs[mid] <= j)
      lo = mid;
    else
      hi = mid;
  }
  owner = lo;
  t = nth_set_bit(L.masks[lo], j - L.offs[lo]);
}

// Publishes every thread's w in L; returns the item count.  Ends with a barrier.
template <int Q>
__device__ __forceinline__ uint32_t publish_items(ItemLds<Q>& L, uint64_t w) {
  uint32_t tot;
  const uint32_t o = block_excl_scan((uint32_t)__popcll(w), &tot);
  L.offs[threadIdx.x] = o;
  L.masks[threadIdx.x] = w;
  if (threadIdx.x == 0) L.offs[BLOCK] = tot;
  if (Q > 0 && tot <= (uint32_t)Q) {
    uint32_t k = o;
    for (uint64_t m = w; m; m &= m - 1) L.q[k++] = (uint16_t)((threadIdx.x << 6) | __builtin_ctzll(m));
  }
  __syncthreads();
  return tot;
}

// Runs f(owner thread, type) over the items of every thread's bit set w.
template <int Q, class F>
__device__ __forceinline__ void run_items(ItemLds<Q>& L, uint64_t w, F&& f) {
  const uint32_t tot = publish_items(L, w);
  for (uint32_t j = threadIdx.x; j < tot; j += BLOCK) {
    int owner, t;
    item_at(L, tot, j, owner, t);
    f(owner, t);
  }
  __syncthreads();
}

// A tile whose pending simplices all wait for the solver gate alone (no verbatim
// sign levels left) finishes in stage 1, as stage 2 would finish it: the LDS tile of
// quantized corners (lds, read no more) holds the owners' references, the accepted
// bits, the corner tables and the item list; the certificate accepts what it can
// (FP64), the solver decides the rest.  Writes the final mask and the tile's count,
// so stage 2 skips the tile; returns false (nothing written) for the other tiles.  (CP2D's
// stage 1 on its det2 tile measured slower with the gate: the batched 1024-job
// pass 32.0 -> 34.4 ms.)
constexpr size_t GATE_LDS = BLOCK * (sizeof(VtxRef) + 8) + sizeof(CornerTab) + sizeof(ItemLds<2048>);
template <int KIND, int D, int M, class T>
__device__ __forceinline__ bool gate_in_stage1(const KParams& P, int64_t blk, int64_t v, const Masks& mk, void* lds,
                                               unsigned long long* prom) {
  if (!__syncthreads_or(mk.sol != 0) || __syncthreads_or(mk.deg != 0)) return false;
  VtxRef* refs = reinterpret_cast<VtxRef*>(lds);
  unsigned long long* slot = reinterpret_cast<unsigned long long*>(refs + BLOCK);
  CornerTab* tab = reinterpret_cast<CornerTab*>(slot + BLOCK);
  ItemLds<2048>& il = *reinterpret_cast<ItemLds<2048>*>(tab + 1);  // the tile's items listed directly
  load_corner_tab<D, M>(P, tab);
  if (mk.sol) refs[threadIdx.x] = make_vtx<D>(P, v).r;
  slot[threadIdx.x] = mk.acc;
  prom[threadIdx.x] = 0;
  uint64_t ws = mk.sol;
  if constexpr (std::is_same<T, double>::value) {
    run_items(il, ws, [&](int owner, int t) {
      T vals[M + 1][M];
      item_values<KIND, D, M, T>(P, *tab, refs[owner], t, vals);
      if (solver_certified<M>(vals))
        atomicOr(&slot[owner], 1ull << t);
      else
        atomicOr(&prom[owner], 1ull << t);
    });
    ws = prom[threadIdx.x];
  }
  if (__syncthreads_or(ws != 0))
    run_items(il, ws, [&](int owner, int t) {
      if (item_solves<KIND, D, M, T>(P, *tab, refs[owner], t)) atomicOr(&slot[owner], 1ull << t);
    });
  const uint64_t acc = slot[threadIdx.x];
  if (v < P.nvert) gptr(P.mask)[v] = acc;
  uint32_t tot;
  block_excl_scan((uint32_t)__popcll(acc), &tot);
  if (threadIdx.x == 0) gptr(P.bsum)[blk] = tot;
  return true;
}

// Stage 1: every vertex's fast classification.  mask = accepted types, pend =
// types that need stage 2 (SoS inside with the solver gate still to run, or the
// verbatim sign levels), degm = the verbatim ones among them.
template <int KIND, int D, int M, class T>
__device__ __forceinline__ void classify_tile(const KParams& P, int64_t blk) {
  constexpr int TBYTES = M == 3   ? Tile3<D>::FLOATS * 4
                         : M == 1 ? (Tile3<D>::NSEG * Tile3<D>::SEG + 15) / 16 * 16
                                  : (TILE2<D, M> ? Tile3<D>::NSEG * Tile3<D>::SEG * 16 : 16);
  __shared__ __align__(16) float tile[TBYTES / 4];
  const int64_t v0 = blk * BLOCK;
  const int64_t v = v0 + threadIdx.x;
  if constexpr (M == 3) {
    stage_quantized3<T, D>(P, v0, tile);
    __syncthreads();
  } else if constexpr (TILE2<D, M>) {
    stage_quantized2<KIND, T, D>(P, v0, reinterpret_cast<double*>(tile));
    __syncthreads();
  } else if constexpr (M == 1) {
    stage_sign_classes<KIND, T, D>(P, v0, reinterpret_cast<int8_t*>(tile));
    __syncthreads();
  }
  Masks mk{0, 0, 0, 0};
  if (v < P.nvert) {
    const Vtx x = make_vtx<D>(P, v);
    mk = classify_fast<KIND, D, M, T>(P, x, tile);
  }
  if constexpr (M == 3) {
    // degenerate tets of vertices whose corners are all below 2^20: the verbatim
    // sign levels on the exact quantized values of the tile (no gathers)
    __shared__ ItemLds<0> il;
    __shared__ unsigned long long prom[BLOCK];
    __shared__ unsigned long long hard[BLOCK];
    __shared__ uint32_t pct[64];
    if (threadIdx.x < 64) pct[threadIdx.x] = packed_chain<D, 3>(threadIdx.x);
    prom[threadIdx.x] = 0;
    hard[threadIdx.x] = 0;
    auto corners = [&](int owner, int t, double (*X)[3]) {
      const uint32_t pc = pct[t];
      double mx = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int s = k == 0 ? 0 : (int)((pc >> (4 * (k - 1))) & 0xF);
        const float* src = tile + ((s >> 1) * Tile3<D>::SEG + owner + (s & 1)) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          X[k][c] = (double)src[c];
          mx = fmax(mx, fabs(X[k][c]));
        }
      }
      return mx;
    };
    // First pass, branch-free: a tet whose twelve quantized values are all zero is
    // inside (every level of the four det3 is zero down to the last, constant one,
    // the same for all four since the corners are in ascending id order; regions
    // that round to zero hold many such tets), and most of the others have one zero
    // vertex (inside3_one_zero).  The rest take the full level cascade in a second
    // pass, spread over the lanes again, so that its branches cost only its items.
    run_items(il, mk.tdeg, [&](int owner, int t) {
      double X[4][3];
      const double mx = corners(owner, t, X);
      const int r = mx == 0.0 ? 1 : inside3_one_zero(X, mx);
      if (r > 0)
        atomicOr(&prom[owner], 1ull << t);
      else if (r < 0)
        atomicOr(&hard[owner], 1ull << t);
    });
    const uint64_t hw = hard[threadIdx.x];
    if (__syncthreads_or(hw != 0))
      run_items(il, hw, [&](int owner, int t) {
        double X[4][3];
        const double mx = corners(owner, t, X);
        if (inside3_exact_d(X, mx)) atomicOr(&prom[owner], 1ull << t);
      });
    mk.sol |= prom[threadIdx.x];
    if constexpr (KIND == KIND_CP) {
      static_assert(GATE_LDS <= sizeof(tile) && sizeof(CornerTab) % 8 == 0, "gate state does not fit");
      if (gate_in_stage1<KIND, D, M, T>(P, blk, v, mk, tile, prom)) return;
    }
  }
  // A tile that leaves nothing to stage 2 has its node count now: it goes to bsum,
  // and stage 2 skips the tile.  Other tiles get a TILE_PENDING sentinel until
  // stage 2 writes their count; when a pending tile has no accepted and no verbatim
  // bits (the usual critical-point tile), mask and degm would be zero and only pend
  // is written (stage 2 writes the final mask).
  const bool pending = __syncthreads_or((mk.sol | mk.deg) != 0);
  const bool sol_only = pending && !__syncthreads_or((mk.acc | mk.deg) != 0);
  if (v < P.nvert) {
    if (sol_only) {
      gptr(P.pend)[v] = mk.sol;
    } else {
      gptr(P.mask)[v] = mk.acc;
      if (pending) {
        gptr(P.pend)[v] = mk.sol | mk.deg;
        gptr(P.degm)[v] = mk.deg;
      }
    }
  }
  if (!pending) {
    uint32_t tot;
    block_excl_scan((uint32_t)__popcll(mk.acc), &tot);
    if (threadIdx.x == 0) gptr(P.bsum)[blk] = tot;
  } else if (threadIdx.x == 0) {
    gptr(P.bsum)[blk] = sol_only ? TILE_PENDING_SOL : TILE_PENDING;
  }
}

// Stage 2: the pending simplices of a tile, spread over the workgroup's lanes by
// run_items so that every lane does useful work; then the final mask and the
// tile's node count.

// Phase A runs the verbatim sign tests of the degenerate items; for critical
// points the ones found inside join the solver items of phase B (B1: the solver
// certificate for FP64 values, B2: the solver for what B1 leaves).  Each phase
// keeps a wave on one code path.
template <int KIND, int D, int M, class T>
__device__ __forceinline__ void stage2_tile(const KParams& P, int64_t blk) {
  // a tile finished by stage 1 has its count in bsum already (one uniform load)
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(gptr(P.bsum)[blk]);
  if (b0 < TILE_PENDING_SOL) return;
  const bool full = b0 == TILE_PENDING;  // else only pend was written
  __shared__ ItemLds<2048> il;
  __shared__ unsigned long long slot[BLOCK];
  __shared__ unsigned long long prom[BLOCK];
  __shared__ VtxRef refs[BLOCK];
  __shared__ CornerTab tab;
  load_corner_tab<D, M>(P, &tab);
  const int64_t v = blk * BLOCK + threadIdx.x;
  uint64_t acc = 0, w = 0, deg = 0;
  if (v < P.nvert) {
    w = gptr(P.pend)[v];
    if (full) {
      acc = gptr(P.mask)[v];
      deg = gptr(P.degm)[v];
    }
  }
  if (__syncthreads_or(w != 0)) {
    if (w) refs[threadIdx.x] = make_vtx<D>(P, v).r;
    slot[threadIdx.x] = acc;
    prom[threadIdx.x] = 0;
    if constexpr (M >= 2) {
      if (__syncthreads_or((w & deg) != 0))  // phase A only when the tile has such items
        run_items(il, w & deg, [&](int owner, int t) {
          if (item_inside<KIND, D, M, T>(P, tab, refs[owner], t)) {
            if constexpr (KIND == KIND_FIBER)
              atomicOr(&slot[owner], 1ull << t);
            else
              atomicOr(&prom[owner], 1ull << t);
          }
        });
    }
    if constexpr (KIND != KIND_FIBER) {
      uint64_t ws = (w & ~deg) | prom[threadIdx.x];
      if constexpr (KIND == KIND_CP && M >= 2 && std::is_same<T, double>::value) {
        // Phase B1: simplices whose solve is certified to succeed are accepted
        // without running the solver; the others go on to phase B2.  Each thread
        // only touches its own prom entry before the first barrier of run_items.
        prom[threadIdx.x] = 0;
        run_items(il, ws, [&](int owner, int t) {
          T vals[M + 1][M];
          item_values<KIND, D, M, T>(P, tab, refs[owner], t, vals);
          if (solver_certified<M>(vals))
            atomicOr(&slot[owner], 1ull << t);
          else
            atomicOr(&prom[owner], 1ull << t);
        });
        ws = prom[threadIdx.x];
      }
      // Phase B2: the solver gate itself
      if (__syncthreads_or(ws != 0))
        run_items(il, ws, [&](int owner, int t) {
          if (item_solves<KIND, D, M, T>(P, tab, refs[owner], t)) atomicOr(&slot[owner], 1ull << t);
        });
    }
    acc = slot[threadIdx.x];
    if (v < P.nvert) gptr(P.mask)[v] = acc;
  }
  uint32_t tot;
  block_excl_scan((uint32_t)__popcll(acc), &tot);
  if (threadIdx.x == 0) gptr(P.bsum)[blk] = tot;
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK, 4) void k_stage2(KParams P) {
  stage2_tile<KIND, D, M, T>(P, blockIdx.x);
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK, 4) void k_classify(KParams P) {
  classify_tile<KIND, D, M, T>(P, blockIdx.x);
}

// exclusive scan of n uint32 in place (single block of 1024 threads); total at a[n]
static __global__ __launch_bounds__(1024) void k_scan_small(uint32_t* a, int64_t n) {
  __shared__ uint32_t part[1024];
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = threadIdx.x * per;
  const int64_t e = b + per < n ? b + per : n;
  uint32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += a[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += y;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (int64_t i = b; i < e; ++i) {
    const uint32_t x = a[i];
    a[i] = run;
    run += x;
  }
  if (threadIdx.x == 1023) a[n] = part[1023];
}

template <int D>
__device__ __forceinline__ bool is_hc_base(const KParams& P, const int64_t* l) {
  bool hc = true;
#pragma unroll
  for (int j = 0; j < D; ++j) hc = hc && (l[j] + 1 < P.dims[j]);
  return hc;
}

template <int KIND, int D, int M, class T>
__device__ __forceinline__ void offsets_tile(const KParams& P, int64_t blk) {
  if (blk == 0 && threadIdx.x == 0) *P.xcount = 0;
  // A tile without nodes has no cells either (every cell of hypercube v has a node
  // owned by v), and no other tile reads its node offsets (only those of vertices
  // that own an active face), so it only records its cell count, 0.
  if (gptr(P.bsum)[blk + 1] == gptr(P.bsum)[blk]) {
    if (threadIdx.x == 0) gptr(P.ebsum)[blk] = 0;
    return;
  }
  const int64_t v = blk * BLOCK + threadIdx.x;
  uint64_t m = 0;
  uint32_t nc = 0;
  if (v < P.nvert) {
    m = gptr(P.mask)[v];
    Vtx x;
    x.v = v;
    decompose<D>(P, v, x.l);
    if (is_hc_base<D>(P, x.l)) nc = cells_of<KIND, D, M>(P, x, 0, (uint32_t*)nullptr);
    if constexpr (CellKind<KIND, D, M>::verts > 0) gptr(P.ccount)[v] = (uint16_t)nc;
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan((uint32_t)__popcll(m), &tot);
  if (v < P.nvert) gptr(P.node_off)[v] = gptr(P.bsum)[blk] + ex;
  block_excl_scan(nc, &tot);
  if (threadIdx.x == 0) gptr(P.ebsum)[blk] = tot;
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void k_offsets(KParams P) {
  offsets_tile<KIND, D, M, T>(P, blockIdx.x);
}

// Lexicographic insertion sort of n cells of NV indices (small n).
template <int NV, class Ptr>
__device__ __forceinline__ void sort_cells(Ptr seg, uint32_t n) {
  for (uint32_t i = 1; i < n; ++i) {
    uint32_t cur[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) cur[q] = seg[i * NV + q];
    int j = (int)i - 1;
    while (j >= 0) {
      bool gt = false, eq = true;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const uint32_t a = seg[j * NV + q];
        if (eq && a != cur[q]) {
          gt = a > cur[q];
          eq = false;
        }
      }
      if (!gt) break;
#pragma unroll
      for (int q = 0; q < NV; ++q) seg[(j + 1) * NV + q] = seg[j * NV + q];
      j--;
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) seg[(j + 1) * NV + q] = cur[q];
  }
}

// Marching cells as workgroup items (contour, d = 4 / 3).  Near a level set only a
// few of a workgroup's vertices have cells, and each of them only in a few of its
// pentatopes (tetrahedra for d = 3); looping over every pentatope of every vertex
// runs both cell shapes for the whole wave whenever one lane has them.  Instead
// every vertex publishes the cells it has (march_masks) and the lanes take one
// (vertex, pentatope) item each (marching_item).  The cells are those of cells_of's
// marching branch; their order inside a hypercube is set afterwards by sorting.

// Pentatope p's cell vertices (chain offsets), four bits each.
template <int D>
struct MarchPack {
  static constexpr uint32_t of(int p) {
    uint32_t r = 0;
    for (int q = 0; q <= D; ++q) r |= (uint32_t)CF<D, 1>.cell[p][q] << (4 * q);
    return r;
  }
  // the base vertex's edge types whose nodes decide pentatope p's shape
  static constexpr uint64_t rel(int p) {
    uint64_t r = 0;
    for (int q = 1; q <= D; ++q) r |= 1ull << (CF<D, 1>.cell[p][q] - 1);
    return r;
  }
};

// act: pentatopes with cells; big: those with the larger shape (three cells for
// d = 4, two for d = 3).  mk0: the base vertex's node mask.
template <int D, int CK>
__device__ __forceinline__ void march_masks(uint64_t mk0, uint32_t* act, uint32_t* big) {
  constexpr int NPC = CellFaceTable<D, 1>::NP;
  uint32_t a = 0, b = 0;
  sfor<0, NPC>([&](auto pp) {
    constexpr int p = decltype(pp)::value;
    const int g1 = __popcll(mk0 & MarchPack<D>::rel(p));
    if (g1 != 0) a |= 1u << p;
    if (CK == 3 ? (g1 == 2 || g1 == 3) : g1 == 2) b |= 1u << p;
  });
  *act = a;
  *big = b;
}

// The cells of marching pentatope / tetrahedron p of the hypercube based at local
// vertex u, each sorted, written to dst (LDS or global memory; 1 or 3 cells for
// d = 4, 1 or 2 for d = 3).
// Edge (i, j) of the cell (positions i < j) carries the node of type
// c_i ^ c_j - 1 owned by vertex u + c_i; vertex pairs are compile-time, so no
// register array is indexed at run time.
template <int D, int CK, class Dst>
__device__ __forceinline__ void marching_item(const KParams& P, const CornerTab& tab, int64_t u, int p, Dst dst) {
  constexpr int NQ = D + 1;
  constexpr int NPC = CellFaceTable<D, 1>::NP;
  constexpr int NV = CK == 3 ? 4 : 3;
  uint32_t pk = 0;
  sfor<0, NPC>([&](auto pp) {
    constexpr int pi = decltype(pp)::value;
    if (p == pi) pk = MarchPack<D>::of(pi);
  });
  int ce[NQ];
  uint64_t mk[NQ - 1];
  uint32_t no[NQ - 1];
#pragma unroll
  for (int q = 0; q < NQ; ++q) ce[q] = (int)((pk >> (4 * q)) & 15u);
#pragma unroll
  for (int q = 0; q < NQ - 1; ++q) {  // the last vertex owns no edge of the cell
    const int64_t w = u + tab.loff[ce[q]];
    mk[q] = gptr(P.mask)[w];
    no[q] = gptr(P.node_off)[w];
  }
  int rel[NQ];
  int g1 = 0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    rel[q] = q == 0 ? 0 : (int)((mk[0] >> (ce[q] - 1)) & 1);
    g1 += rel[q];
  }
  auto E = [&](int i, int j) -> uint32_t {  // i < j, both compile-time after unrolling
    return no[i] + (uint32_t)__popcll(mk[i] & ((1ull << ((ce[i] ^ ce[j]) - 1)) - 1ull));
  };
  auto put = [&](int k, uint32_t* c) {
    sort_small<NV>(c);
#pragma unroll
    for (int i = 0; i < NV; ++i) dst[k * NV + i] = c[i];
  };
  const bool single = g1 == 1 || g1 == NQ - 1;
  if (single || (CK == 4 && g1 == 2)) {
    // the D active edges (one vertex alone: its edges; d = 3, two and two: the four
    // edges between the pairs), in any order
    uint32_t c[4] = {0, 0, 0, 0};
    int k = 0;
    sfor<0, NQ>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      sfor<i + 1, NQ>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        if (rel[i] != rel[j]) {
          const uint32_t e = E(i, j);
#pragma unroll
          for (int s = 0; s < 4; ++s)
            if (k == s) c[s] = e;
          k++;
        }
      });
    });
    if (single) {
      put(0, c);
    } else {
      sort_small<4>(c);
      uint32_t c0[3] = {c[0], c[1], c[2]};
      uint32_t c1[3] = {c[0], c[2], c[3]};
      put(0, c0);
      put(1, c1);
    }
  } else if constexpr (CK == 3) {
    // two against three: T[3 * (rank of a in A) + (rank of b in B)] = node of edge
    // (a, b), A the pair, ranks by position; the prism splits along its staircase
    const int relA = g1 == 2 ? 1 : 0;
    int rA[NQ], rB[NQ];
    int na = 0, nb = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      rA[q] = na;
      rB[q] = nb;
      if (rel[q] == relA)
        na++;
      else
        nb++;
    }
    uint32_t T[6] = {0, 0, 0, 0, 0, 0};
    sfor<0, NQ>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      sfor<i + 1, NQ>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        if (rel[i] != rel[j]) {
          const int slot = rel[i] == relA ? 3 * rA[i] + rB[j] : 3 * rA[j] + rB[i];
          const uint32_t e = E(i, j);
#pragma unroll
          for (int s = 0; s < 6; ++s)
            if (slot == s) T[s] = e;
        }
      });
    });
    uint32_t c0[4] = {T[0], T[1], T[2], T[5]};
    uint32_t c1[4] = {T[0], T[1], T[4], T[5]};
    uint32_t c2[4] = {T[0], T[3], T[4], T[5]};
    put(0, c0);
    put(1, c1);
    put(2, c2);
  }
}

// Position of staged cell c inside its vertex's segment [s, e) of the staging
// array: the number of the segment's cells ordered before it (lexicographic, equal
// cells in staging order).  One thread per cell, so a vertex with many cells costs
// its segment length per cell instead of a quadratic sort in one lane; with step > 1
// the count over cells s, s + step, ... (the caller sums the lanes' parts).
template <int NV>
__device__ __forceinline__ uint32_t cell_rank(const uint32_t* stage, uint32_t s, uint32_t e, uint32_t c,
                                              uint32_t step = 1) {
  uint32_t r = 0;
  if constexpr (NV == 2) {
    const uint64_t kc = ((uint64_t)stage[2 * c] << 32) | stage[2 * c + 1];
    for (uint32_t j = s; j < e; ++j) {
      const uint64_t kj = ((uint64_t)stage[2 * j] << 32) | stage[2 * j + 1];
      r += (kj < kc || (kj == kc && j < c)) ? 1u : 0u;
    }
  } else {
    // two 64-bit keys: the first two indices, then the rest
    auto hi = [&](uint32_t j) { return ((uint64_t)stage[j * NV] << 32) | stage[j * NV + 1]; };
    auto lo = [&](uint32_t j) {
      return NV == 4 ? ((uint64_t)stage[j * NV + 2] << 32) | stage[j * NV + NV - 1] : (uint64_t)stage[j * NV + 2];
    };
    const uint64_t ch = hi(c), cl = lo(c);
#pragma unroll 4
    for (uint32_t j = s; j < e; j += step) {
      const uint64_t jh = hi(j), jl = lo(j);
      r += (jh < ch || (jh == ch && (jl < cl || (jl == cl && j < c)))) ? 1u : 0u;
    }
  }
  return r;
}

// Emit: the nodes of a workgroup occupy one contiguous output range in canonical
// order (vertex, then type), so they are queued in LDS in that order and node j
// of the workgroup is solved and written by thread j mod 256 (coalesced, load
// balanced).  Cells are staged in LDS, each vertex's segment sorted, and copied
// out contiguously.
constexpr int EMIT_NQ = 6144;          // queued nodes per workgroup
constexpr int EMIT_LUF = EMIT_NQ / 2;   // nodes of the workgroup union-find (u32 in nq)

// Union-find in LDS over one workgroup's nodes (hook the larger root under the
// smaller, so roots are minima).
__device__ __forceinline__ uint32_t lds_find(uint32_t* p, uint32_t x) {
  while (true) {
    const uint32_t y = __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (y == x) return x;
    const uint32_t z = __hip_atomic_load(p + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (z != y) __hip_atomic_store(p + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = y;
  }
}
__device__ __forceinline__ void lds_unite(uint32_t* p, uint32_t a, uint32_t b) {
  while (true) {
    a = lds_find(p, a);
    b = lds_find(p, b);
    if (a == b) return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t expected = b;
    if (__hip_atomic_compare_exchange_strong(p + b, &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
  }
}
constexpr int EMIT_STAGE_WORDS = 4096;  // staged cell indices per workgroup

// Workgroup union-find of cells with three or four nodes (ntot <= EMIT_LUF nodes,
// lpar in the node queue's memory; xh: EMIT_XH slots of free LDS; the cells in LDS
// or, for a workgroup whose cells overflow the staging array, in global memory).
// Each cell's nodes inside the workgroup are united in LDS through its first such
// node; every node outside it is paired with that node's root (with the cell's
// first node when none is inside) for the global union-find.  A node lies in many
// cells, so the pairs repeat: an LDS hash set keeps one of each (a pair that finds
// no slot is kept, a repeated union being harmless).  Ends with the pairs written.
constexpr int EMIT_XH = 1024;
template <int NV, class Src>
__device__ __forceinline__ void emit_link_cells(const KParams& P, uint16_t* nq, void* free_lds, Src cl,
                                                uint32_t nbase, uint32_t ntot, uint32_t ctot) {
  static_assert(NV >= 3 && NV <= 4, "cells of three or four nodes");
  constexpr unsigned long long EMPTY = ~0ull;
  uint32_t* lpar = reinterpret_cast<uint32_t*>(nq);
  unsigned long long* xh = reinterpret_cast<unsigned long long*>(free_lds);
  __syncthreads();  // the caller's last reads of that LDS (the cell owners) are done
  for (uint32_t i = threadIdx.x; i < ntot; i += BLOCK) lpar[i] = i;
  for (uint32_t i = threadIdx.x; i < (uint32_t)EMIT_XH; i += BLOCK) xh[i] = EMPTY;
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < ctot; c += BLOCK) {
    uint32_t f = ~0u;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const uint32_t n = cl[c * NV + q] - nbase;
      if (n < ntot) {
        if (f == ~0u)
          f = n;
        else
          lds_unite(lpar, f, n);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ntot; i += BLOCK) gptr(P.parent)[nbase + i] = nbase + lds_find(lpar, i);
  // the pairs of cell c: anchor (root of its first inside node, else its first node)
  // with each outside node other than the anchor
  auto anchor_of = [&](uint32_t c, bool* inside) {
    uint32_t f = ~0u;
#pragma unroll
    for (int q = NV - 1; q >= 0; --q) {
      const uint32_t n = cl[c * NV + q] - nbase;
      if (n < ntot) f = n;
    }
    *inside = f != ~0u;
    return f != ~0u ? nbase + lds_find(lpar, f) : (uint32_t)cl[c * NV];
  };
  // rounds of CPR cells per thread: bit (round cell) * NV + q of `own` marks the
  // pairs this thread keeps
  constexpr uint32_t CPR = 32 / NV;
  __shared__ uint32_t xbase;
  for (uint32_t base = 0; base < ctot; base += CPR * BLOCK) {
    const uint32_t end = min(ctot, base + CPR * BLOCK);
    uint32_t own = 0, nx = 0;
    int bit = 0;
    for (uint32_t c = base + threadIdx.x; c < end; c += BLOCK, bit += NV) {
      bool inside;
      const uint32_t anc = anchor_of(c, &inside);
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const uint32_t g = cl[c * NV + q];
        if (g - nbase < ntot || (!inside && q == 0)) continue;
        const unsigned long long key =
            ((unsigned long long)min(anc, g) << 32) | (unsigned long long)max(anc, g);
        const uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 54);  // 10 bits: EMIT_XH slots
        bool keep = true;
#pragma unroll
        for (int probe = 0; probe < 8; ++probe) {
          const unsigned long long old = atomicCAS(&xh[(h + probe) & (EMIT_XH - 1)], EMPTY, key);
          if (old == EMPTY) break;
          if (old == key) {
            keep = false;
            break;
          }
        }
        if (keep) {
          own |= 1u << (bit + q);
          nx++;
        }
      }
    }
    uint32_t xtot;
    uint32_t xo = block_excl_scan(nx, &xtot);
    if (threadIdx.x == 0) xbase = xtot ? atomicAdd(P.xcount, xtot) : 0;
    __syncthreads();
    xo += xbase;
    bit = 0;
    for (uint32_t c = base + threadIdx.x; c < end && own; c += BLOCK, bit += NV) {
      if (((own >> bit) & ((1u << NV) - 1u)) == 0) continue;
      bool inside;
      const uint32_t anc = anchor_of(c, &inside);
#pragma unroll
      for (int q = 0; q < NV; ++q)
        if ((own >> (bit + q)) & 1u) {
          gptr(P.xedges)[2 * (uint64_t)xo] = anc;
          gptr(P.xedges)[2 * (uint64_t)xo + 1] = cl[c * NV + q];
          xo++;
        }
    }
  }
}

template <int KIND, int D, int M, class T>
__device__ __forceinline__ void emit_tile(const KParams& P, int64_t blk) {
  constexpr int NTY = Chains<D, M>::N;
  constexpr int NV = CellKind<KIND, D, M>::verts;
  // a tile without nodes has no cells (k_offsets left its counts unwritten)
  if (gptr(P.bsum)[blk + 1] == gptr(P.bsum)[blk]) return;
  __shared__ __align__(16) uint16_t nq[EMIT_NQ];  // node queue, marching items, then the union-find
  __shared__ VtxRef refs[BLOCK];
  __shared__ uint32_t stage[EMIT_STAGE_WORDS];
  __shared__ CornerTab tab;
  load_corner_tab<D, M>(P, &tab);
  __syncthreads();
  const int64_t v = blk * BLOCK + threadIdx.x;
  Vtx x;
  uint64_t m = 0;
  bool hc = false;
  // every per-vertex and per-tile input of the tile is requested here, before the
  // first barrier (one memory round trip instead of one per phase)
  uint32_t pre_off = 0, pre_nc = 0;
  if (v < P.nvert) {
    x = make_vtx<D>(P, v);
    m = gptr(P.mask)[v];
    pre_off = gptr(P.node_off)[v];
    if constexpr (NV > 0) pre_nc = gptr(P.ccount)[v];
    hc = is_hc_base<D>(P, x.l);
    refs[threadIdx.x] = x.r;
  }
  const uint64_t pre_cbase = NV > 0 ? gptr(P.ebsum)[blk] : 0;
  // ---- nodes
  const uint32_t nbase = gptr(P.bsum)[blk];
  const uint32_t ntot = gptr(P.bsum)[blk + 1] - nbase;
  auto lam_out = gptr(static_cast<T*>(P.lambda));
  auto emit_from = [&](uint32_t pos, const VtxRef& r, int t, const T (*vals)[M]) {
    T lam[M + 1];
    float sc;
    node_solve<KIND, D, M, T>(P, tab, r, t, vals, lam, &sc);
    gptr(P.keys)[pos] = r.gid * (uint64_t)NTY + (uint64_t)t;
#pragma unroll
    for (int i = 0; i <= M; ++i) lam_out[(uint64_t)pos * (M + 1) + i] = lam[i];
    if (P.scalar_out) gptr(P.scalar_out)[pos] = sc;
  };
  if (ntot <= P.nq_cap) {
    if (m) {
      uint32_t pos = pre_off - nbase;
      for (uint64_t mm = m; mm; mm &= mm - 1) nq[pos++] = (uint16_t)(threadIdx.x | (__builtin_ctzll(mm) << 8));
    }
    __syncthreads();
    // software pipelined: the next node's gathers are in flight while this one is solved
    uint32_t i = threadIdx.x;
    if (i < ntot) {
      uint32_t it = nq[i];
      T cur[M + 1][M];
      node_values<KIND, D, M, T>(P, tab, refs[it & 0xFF], (int)(it >> 8), cur);
      while (true) {
        const uint32_t j = i + BLOCK;
        uint32_t jt = 0;
        T nxt[M + 1][M];
        if (j < ntot) {
          jt = nq[j];
          node_values<KIND, D, M, T>(P, tab, refs[jt & 0xFF], (int)(jt >> 8), nxt);
        }
        emit_from(nbase + i, refs[it & 0xFF], (int)(it >> 8), cur);
        if (j >= ntot) break;
        i = j;
        it = jt;
#pragma unroll
        for (int a = 0; a <= M; ++a)
#pragma unroll
          for (int k = 0; k < M; ++k) cur[a][k] = nxt[a][k];
      }
    }
  } else if (m) {  // more nodes than the queue holds: each thread writes its own
    uint32_t pos = pre_off;
    for (uint64_t mm = m; mm; mm &= mm - 1) {
      const int t = __builtin_ctzll(mm);
      T vals[M + 1][M];
      node_values<KIND, D, M, T>(P, tab, x.r, t, vals);
      emit_from(pos++, x.r, t, vals);
    }
  }
  bool linked = false;
  // ---- cells
  if constexpr (NV > 0) {
    // the count k_offsets made for this hypercube
    const uint32_t nc = hc ? pre_nc : 0u;
    uint32_t ctot;
    const uint32_t ex = block_excl_scan(nc, &ctot);
    const uint64_t cbase = pre_cbase;
    __shared__ uint32_t segx[BLOCK + 1];  // offset of every thread's segment of cells
    constexpr int CKV = CellKind<KIND, D, M>::value;
    // marching cells as (vertex, pentatope) items written at cells + k * NV (k: the
    // cell's offset in the workgroup), the owner thread of each cell recorded in
    // cown when given; the node queue and the vertex references of the node phase
    // are free.  Ends with a barrier.
    auto march = [&](auto cells, uint8_t* cown) {
      static_assert(sizeof(ItemLds<0>) <= sizeof(nq) && sizeof(refs) >= 2 * BLOCK * 4, "marching items do not fit");
      uint32_t act = 0, big = 0;
      if (nc) march_masks<D, CKV>(m, &act, &big);
      uint32_t* actl = reinterpret_cast<uint32_t*>(refs);
      uint32_t* bigl = actl + BLOCK;
      actl[threadIdx.x] = act;
      bigl[threadIdx.x] = big;
      run_items(*reinterpret_cast<ItemLds<0>*>(nq), (uint64_t)act, [&](int owner, int p) {
        const uint32_t below = (1u << p) - 1u;
        const uint32_t k = segx[owner] + (uint32_t)__popc(actl[owner] & below) +
                           (CKV == 3 ? 2u : 1u) * (uint32_t)__popc(bigl[owner] & below);
        marching_item<D, CKV>(P, tab, blk * BLOCK + owner, p, cells + k * NV);
        if (cown) {
          const uint32_t ncell = ((bigl[owner] >> p) & 1u) ? (CKV == 3 ? 3u : 2u) : 1u;
          for (uint32_t i = 0; i < ncell; ++i) cown[k + i] = (uint8_t)owner;
        }
      });
    };
    segx[threadIdx.x] = ex;
    if (threadIdx.x == 0) segx[BLOCK] = ctot;
    if (ctot * NV <= P.stage_cap) {
      // The thread whose segment holds each cell (at most EMIT_STAGE_WORDS / 2
      // cells), in the LDS of the node phase's vertex references, which are no
      // longer read, where its writes are cheap: star edges (two or three per
      // vertex) by their vertex's thread, marching cells (one to three per item) by
      // the item's lane.  Star-fan cells come many per vertex, where one thread's
      // serial byte stores cost more than a binary search of segx per cell (so did
      // marching cells written per vertex: contour emit 45.5 -> 66.7 ms).
      constexpr bool OWNERS = NV == 2 || CKV == 3 || CKV == 4;
      constexpr int OWN_AT = (CKV == 3 || CKV == 4) ? 2 * BLOCK * 4 : 0;  // after the items' masks
      static_assert(sizeof(refs) >= OWN_AT + EMIT_STAGE_WORDS / 2, "cell owners do not fit");
      uint8_t* cown = reinterpret_cast<uint8_t*>(refs) + OWN_AT;
      if constexpr (CKV == 3 || CKV == 4) {
        march(stage, cown);
      } else {
        if (nc) {
          cells_of<KIND, D, M>(P, x, 1, stage + ex * NV);
          if constexpr (OWNERS)
            for (uint32_t c = ex; c < ex + nc; ++c) cown[c] = (uint8_t)threadIdx.x;
        }
        __syncthreads();
      }
      // each cell goes straight to its sorted place in the workgroup's output range
      auto dst = gptr(P.cells) + cbase * NV;
      // Cells of three or four nodes come few per workgroup but many per vertex
      // (up to 72 per hypercube for marching cells): four consecutive lanes count
      // each one's rank over every fourth cell of its segment, so that the LDS
      // reads of a long segment are spread over more lanes.  Lane groups are whole
      // (ctot * G and BLOCK are multiples of G), so the shuffles see all four.
      constexpr uint32_t G = NV >= 3 ? 4 : 1;
      for (uint32_t w = threadIdx.x; w < ctot * G; w += BLOCK) {
        const uint32_t c = w / G, part = w % G;
        uint32_t lo;
        if constexpr (OWNERS) {
          lo = cown[c];
        } else {
          uint32_t hi = BLOCK;  // segx[lo] <= c < segx[hi]
          lo = 0;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (segx[mid] <= c)
              lo = mid;
            else
              hi = mid;
          }
        }
        const uint32_t sb = segx[lo];
        uint32_t r = cell_rank<NV>(stage, sb + part, segx[lo + 1], c, G);
        if constexpr (G > 1) {
          r += __shfl_xor(r, 1);
          r += __shfl_xor(r, 2);
        }
        if (part == 0) {
          const uint32_t pos = sb + r;
#pragma unroll
          for (int q = 0; q < NV; ++q) dst[pos * NV + q] = stage[c * NV + q];
        }
      }
      // Union the cells whose nodes all belong to this workgroup in LDS and start
      // the global union-find from those roots (roots stay component minima).
      if (NV >= 3 && ntot <= EMIT_LUF) {
        static_assert(sizeof(refs) >= EMIT_XH * 8, "the pair set does not fit");
        emit_link_cells<NV < 3 ? 3 : NV>(P, nq, refs, static_cast<const uint32_t*>(stage), nbase, ntot, ctot);
        linked = true;
      } else if (ntot <= EMIT_LUF) {
        uint32_t* lpar = reinterpret_cast<uint32_t*>(nq);  // the node queue is done
        for (uint32_t i = threadIdx.x; i < ntot; i += BLOCK) lpar[i] = i;
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < ctot; c += BLOCK) {
          const uint32_t a = stage[c * NV] - nbase;
          if (a >= ntot) continue;
#pragma unroll
          for (int q = 1; q < NV; ++q) {
            const uint32_t b = stage[c * NV + q] - nbase;
            if (b < ntot) lds_unite(lpar, a, b);
          }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < ntot; i += BLOCK) gptr(P.parent)[nbase + i] = nbase + lds_find(lpar, i);
        // edges leaving the workgroup, from the local root of their first node
        uint32_t nx = 0;
        for (uint32_t c = threadIdx.x; c < ctot; c += BLOCK) {
          const uint32_t a = stage[c * NV] - nbase;
#pragma unroll
          for (int q = 1; q < NV; ++q) nx += (a >= ntot || stage[c * NV + q] - nbase >= ntot) ? 1u : 0u;
        }
        uint32_t xtot;
        uint32_t xo = block_excl_scan(nx, &xtot);
        __shared__ uint32_t xbase;
        if (threadIdx.x == 0) xbase = xtot ? atomicAdd(P.xcount, xtot) : 0;
        __syncthreads();
        xo += xbase;
        for (uint32_t c = threadIdx.x; c < ctot; c += BLOCK) {
          const uint32_t a = stage[c * NV] - nbase;
          const uint32_t ra = a < ntot ? nbase + lds_find(lpar, a) : stage[c * NV];
#pragma unroll
          for (int q = 1; q < NV; ++q) {
            const uint32_t b = stage[c * NV + q];
            if (a >= ntot || b - nbase >= ntot) {
              gptr(P.xedges)[2 * (uint64_t)xo] = ra;
              gptr(P.xedges)[2 * (uint64_t)xo + 1] = b;
              xo++;
            }
          }
        }
        linked = true;
      }
    } else {
      auto seg = gptr(P.cells) + (cbase + ex) * NV;
      if constexpr (CKV == 3 || CKV == 4) {
        // straight to the output range, then each vertex's thread sorts its segment
        // (the items' writes are visible to the workgroup after their barrier)
        march(gptr(P.cells) + cbase * NV, static_cast<uint8_t*>(nullptr));
        if (nc) sort_cells<NV>(seg, nc);
      } else if (nc) {
        cells_of<KIND, D, M>(P, x, 1, seg);
        sort_cells<NV>(seg, nc);
      }
      if (NV >= 3 && ntot <= EMIT_LUF) {
        // more cells than the staging array holds: the same union-find over the
        // cells in global memory (each thread's own writes are visible to the
        // workgroup after the barrier)
        static_assert(sizeof(refs) >= EMIT_XH * 8, "the pair set does not fit");
        __syncthreads();
        const FTK2H_AS1 uint32_t* src = gptr(P.cells) + cbase * NV;
        emit_link_cells<NV < 3 ? 3 : NV>(P, nq, refs, src, nbase, ntot, ctot);
        linked = true;
      }
    }
    if (!linked) {
      // every edge of this workgroup goes to the global union-find
      __syncthreads();  // cells written to global memory by their threads
      const uint32_t ne = ctot * (NV - 1);
      __shared__ uint32_t xbase2;
      if (threadIdx.x == 0) xbase2 = ne ? atomicAdd(P.xcount, ne) : 0;
      __syncthreads();
      auto src = gptr(P.cells) + cbase * NV;
      for (uint32_t e = threadIdx.x; e < ne; e += BLOCK) {
        const uint32_t c = e / (NV - 1), q = 1 + e % (NV - 1);
        gptr(P.xedges)[2 * (uint64_t)(xbase2 + e)] = src[c * NV];
        gptr(P.xedges)[2 * (uint64_t)(xbase2 + e) + 1] = src[c * NV + q];
      }
    }
  }
  if (!linked)
    for (uint32_t i = threadIdx.x; i < ntot; i += BLOCK) gptr(P.parent)[nbase + i] = nbase + i;
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void k_emit(KParams P) {
  emit_tile<KIND, D, M, T>(P, blockIdx.x);
}

// union-find hooking (larger root under smaller).  COMPRESS: finds shorten the
// paths they walk (path splitting).  Cross-tile linking of star edges runs without
// it: their forest is shallow after the workgroup unions, and the stores cost more
// than they save (256^3 x 4 CP3D step: linking 1.95 -> 1.79 ms, the final compress
// pass 0.79 -> 0.83 ms); the denser marching cells need it (contour 384^3 x 4
// linking 9.0 -> 11.6 ms without), and so do unions straight over a slab's cells.
__device__ __forceinline__ uint32_t uf_load(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool COMPRESS>
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  uint32_t p = uf_load(parent + x);
  while (p != x) {
    const uint32_t gp = uf_load(parent + p);
    if (COMPRESS && gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = gp;
  }
  return x;
}
template <bool COMPRESS>
__device__ __forceinline__ void uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
  while (true) {
    a = uf_find<COMPRESS>(parent, a);
    b = uf_find<COMPRESS>(parent, b);
    if (a == b) return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t expected = b;
    if (__hip_atomic_compare_exchange_strong(parent + b, &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
    b = expected;
  }
}

// Grid-stride over the *count pairs (the count is only known on the device, so the
// grid is a fixed persistent size instead of one thread per possible pair).
template <bool COMPRESS>
__global__ __launch_bounds__(BLOCK) void k_link_pairs(uint32_t* parent, const uint32_t* pairs, const uint32_t* count) {
  const uint64_t n = *count;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
    uf_unite<COMPRESS>(parent, pairs[2 * i], pairs[2 * i + 1]);
}
constexpr unsigned LINK_GRID = 256 * 16;  // workgroups of the persistent link kernel

static __global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = v;
}

// track[i] = root of i.  The walk is latency-bound, so each thread follows
// COMPRESS_W independent chains at once (more loads in flight per wave).
constexpr int COMPRESS_W = 4;
static __global__ __launch_bounds__(BLOCK) void k_compress(uint32_t* parent, uint32_t* track, uint64_t n) {
  const uint64_t i0 = (uint64_t)blockIdx.x * BLOCK * COMPRESS_W + threadIdx.x;
  uint32_t x[COMPRESS_W], p[COMPRESS_W];
#pragma unroll
  for (int j = 0; j < COMPRESS_W; ++j) {
    const uint64_t i = i0 + (uint64_t)j * BLOCK;
    x[j] = (uint32_t)i;
    p[j] = i < n ? parent[i] : (uint32_t)i;  // beyond n: already "at its root", never loads
  }
  while (true) {
    bool more = false;
#pragma unroll
    for (int j = 0; j < COMPRESS_W; ++j)
      if (p[j] != x[j]) {
        x[j] = p[j];
        p[j] = parent[x[j]];
        more = true;
      }
    if (!more) break;
  }
#pragma unroll
  for (int j = 0; j < COMPRESS_W; ++j) {
    const uint64_t i = i0 + (uint64_t)j * BLOCK;
    if (i < n) track[i] = x[j];
  }
}
static inline void launch_compress(uint32_t* parent, uint32_t* track, uint64_t n, hipStream_t st) {
  const uint64_t per = (uint64_t)BLOCK * COMPRESS_W;
  if (n) k_compress<<<dim3((unsigned)((n + per - 1) / per)), dim3(BLOCK), 0, st>>>(parent, track, n);
}

}  // namespace ftk2h
