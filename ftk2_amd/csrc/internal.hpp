// Internal declarations shared by the host translation units of libftk2h:
// error reporting, device buffers, context/result objects and the templated
// pipeline driver (instantiated per predicate in inst_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ftk2h.h"
#include "kernels.hpp"
#include "sortcells.hpp"

#include <algorithm>
#include <functional>
#include <map>

namespace ftk2h_internal {
using namespace ftk2h;

std::string& error_slot();

inline int fail(int code, const std::string& msg) {
  error_slot() = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(e_ == hipErrorOutOfMemory ? FTK2H_E_OOM : FTK2H_E_HIP,              \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                 \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return FTK2H_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) return FTK2H_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? FTK2H_E_OOM : FTK2H_E_HIP,
                  std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = n;
    return FTK2H_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

constexpr int NPHASE = 8;

// Cache of device blocks for result arrays: a freed result returns its blocks
// here and the next extraction of a similar size takes them back instead of
// calling hipMalloc/hipFree (hundreds of microseconds each for GB-sized arrays).
// Bounded by `cap` bytes of idle blocks; shared by the context and its results.
struct DevPool {
  int device = 0;
  size_t cap = 0, idle = 0;
  std::mutex m;
  std::multimap<size_t, void*> free_blocks;  // capacity -> block
  std::map<void*, size_t> owned;              // every block this pool handed out
  hipError_t get(void** out, size_t n) {
    {
      std::lock_guard<std::mutex> g(m);
      auto it = free_blocks.lower_bound(n);
      if (it != free_blocks.end() && it->first <= 2 * n + (1u << 20)) {
        *out = it->second;
        idle -= it->first;
        free_blocks.erase(it);
        return hipSuccess;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e == hipErrorOutOfMemory) {  // give the idle blocks back and retry once
      trim(0);
      e = hipMalloc(&p, n ? n : 1);
    }
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(m);
    owned[p] = n ? n : 1;
    *out = p;
    return hipSuccess;
  }
  void put(void* p) {
    if (!p) return;
    std::unique_lock<std::mutex> g(m);
    auto it = owned.find(p);
    if (it == owned.end() || idle + it->second > cap) {
      if (it != owned.end()) owned.erase(it);
      g.unlock();
      free_on_device(p);
      return;
    }
    free_blocks.emplace(it->second, p);
    idle += it->second;
  }
  void trim(size_t keep) {
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> g(m);
      while (idle > keep && !free_blocks.empty()) {
        auto it = std::prev(free_blocks.end());
        idle -= it->first;
        owned.erase(it->second);
        drop.push_back(it->second);
        free_blocks.erase(it);
      }
    }
    for (void* p : drop) free_on_device(p);
  }
  void free_on_device(void* p) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(p);
    (void)hipSetDevice(cur);
  }
  ~DevPool() { trim(0); }
};

}  // namespace ftk2h_internal

struct ftk2h_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  ftk2h_internal::DevBuf mask, pend, degm, ccount, node_off, bsum, ebsum, xedges, xcount, scan_sums, fields[4];
  hipEvent_t ev[ftk2h_internal::NPHASE + 1] = {};
  std::shared_ptr<ftk2h_internal::DevPool> pool;
  ftk2h_internal::DevBuf parent;  // union-find scratch
  std::mutex lock;
};

struct ftk2h_result {
  int device = 0;
  int kind = 0, d = 0, m = 0, dtype = 0;
  uint64_t gdims[4] = {1, 1, 1, 1};
  uint64_t tstride = 0;  // id step of the last axis when extruded (0: product of gdims)
  uint64_t n_nodes = 0;
  int ntypes = 0;
  uint64_t* keys = nullptr;
  void* lambda = nullptr;
  uint32_t* track = nullptr;
  float* scalar = nullptr;
  float scalar_value = 0.f;  // FeatureElement::scalar of every node when `scalar` is null
  float* attrs = nullptr;  // n_nodes * 16 (FeatureElement::attributes), when requested
  uint32_t* types = nullptr;  // n_nodes (FeatureElement::type), when classified
  int n_conn = 0;
  int conn_dim[3] = {0, 0, 0};
  uint64_t conn_cells[3] = {0, 0, 0};
  uint32_t* conn[3] = {nullptr, nullptr, nullptr};
  float phase_ms[ftk2h_internal::NPHASE] = {};
  bool has_scalar = false;
  // Set when the arrays are views into memory shared with other results (batched
  // runs); the memory is released with its last holder.
  std::shared_ptr<void> arena;
  // Device blocks of keys/lambda/track/scalar/conn come from this pool (or, when
  // it is null, from hipMalloc directly).
  std::shared_ptr<ftk2h_internal::DevPool> pool;
  void release(void* p) {
    if (!p) return;
    if (pool) {
      pool->put(p);
      return;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(p);
    (void)hipSetDevice(cur);
  }
  ~ftk2h_result() {
    release(attrs);  // never part of a batch's shared memory
    release(types);
    if (arena) return;
    release(keys);
    release(lambda);
    release(track);
    release(scalar);
    for (auto* c : conn) release(c);
  }
};

namespace ftk2h_internal {

// Output arrays supplied by the caller and work memory that must not be allocated
// (ftk2h_extract_regular_noalloc): the pipeline writes the result there and fails
// with FTK2H_E_OVERFLOW when a capacity is too small.
struct FixedOut {
  uint64_t* keys;
  void* lambda;
  uint32_t* track;
  float* scalar;
  uint32_t* cells;
  uint64_t node_cap, cell_cap;  // caller's capacities (nodes, cells)
  char* tail;                   // scratch for the union-find parents and the cross-workgroup
  uint64_t tail_bytes;          // edges, split once the node and cell counts are known
  uint32_t* scan_sums;
};

struct RunArgs {
  ftk2h_ctx* ctx;
  hipStream_t st;
  KParams P;
  int64_t nb;  // blocks over vertices
  uint64_t node_cap, cell_cap;
  bool sync;
  ftk2h_result* res;
  const FixedOut* fixed = nullptr;  // null: result arrays from the context's pool
};

using RunFn = int (*)(RunArgs&);

// Entry-point helpers (ftk2h.hip).
int check_request(const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred, const ftk2h_field_view* comps,
                  int ncomp, RunFn* fn);
void init_result(ftk2h_result* res, const ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh,
                 const ftk2h_predicate_desc* pred);
int bind_request(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred,
                 const ftk2h_field_view* comps, int ncomp, const ftk2h_field_view* scalar, bool allow_host,
                 hipStream_t st, KParams& P, uint64_t* nvert, bool* has_cells);
// Device address of the fixed scratch array on the current device (noalloc.hip); called
// once per device when a context is created.
int noalloc_init(int device);

// ---- request checks and kernel parameters shared by every entry point ----------

// Number of field components the predicate reads, or -1 (message set) if the
// (kind, codimension) pair is invalid.
inline int components_needed(const ftk2h_predicate_desc* pred) {
  switch (pred->kind) {
    case FTK2H_PRED_CRITICAL_POINT:
      return pred->codimension;
    case FTK2H_PRED_CONTOUR:
      if (pred->codimension != 1) return fail(FTK2H_E_INVALID, "contour codimension must be 1"), -1;
      return 1;
    case FTK2H_PRED_FIBER:
      if (pred->codimension != 2) return fail(FTK2H_E_INVALID, "fiber codimension must be 2"), -1;
      return 2;
    default:
      return fail(FTK2H_E_INVALID, "unknown predicate kind"), -1;
  }
}

// Validates a mesh descriptor against the predicate's dimension d; returns the
// vertex count through nvert and whether the mesh has any hypercube.
inline int check_mesh(const ftk2h_mesh_desc* mesh, int d, uint64_t* nvert, bool* has_cells) {
  if (mesh->ndims != d) return fail(FTK2H_E_INVALID, "mesh ndims differs from the predicate's");
  uint64_t n = 1;
  bool hc = true;
  for (int j = 0; j < d; ++j) {
    if (mesh->local_dims[j] == 0) return fail(FTK2H_E_INVALID, "zero extent");
    if (mesh->global_dims[j] < mesh->local_dims[j] + mesh->offset[j])
      return fail(FTK2H_E_INVALID, "local dims + offset exceed global dims");
    n *= mesh->local_dims[j];
    hc = hc && mesh->local_dims[j] >= 2;
  }
  if (mesh->time_id_stride) {
    uint64_t sp = 1;
    for (int j = 0; j + 1 < d; ++j) sp *= mesh->global_dims[j];
    if (mesh->time_id_stride < sp) return fail(FTK2H_E_INVALID, "time_id_stride below the spatial vertex count");
    if (mesh->global_dims[d - 1] > (1ull << 57) / mesh->time_id_stride)
      return fail(FTK2H_E_INVALID, "vertex ids beyond 2^57");
  }
  *nvert = n;
  *has_cells = hc;
  return FTK2H_OK;
}

// Mesh and predicate fields of P (field pointers, tables and work buffers excluded).
inline void init_params(KParams& P, const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred, uint64_t nvert) {
  std::memset(&P, 0, sizeof(P));
  const int d = mesh->ndims;
  for (int j = 0; j < 4; ++j) {
    P.dims[j] = j < d ? (int64_t)mesh->local_dims[j] : 1;
    P.off[j] = j < d ? (int64_t)mesh->offset[j] : 0;
  }
  int64_t gs = 1;
  for (int j = 0; j < d; ++j) {
    P.gstride[j] = gs;
    gs *= (int64_t)mesh->global_dims[j];
  }
  if (mesh->time_id_stride) P.gstride[d - 1] = (int64_t)mesh->time_id_stride;  // extruded layers
  P.nvert = (int64_t)nvert;
  P.small = nvert < (1ull << 31);
  P.Q = pred->sos_q;
  P.thr[0] = pred->thresholds[0];
  P.thr[1] = pred->thresholds[1];
}

// Host view -> contiguous x-fastest copy of the local grid.
inline void pack_host_view(const ftk2h_field_view& fv, const ftk2h_mesh_desc* mesh, size_t esz, uint64_t nvert,
                           char* dst) {
  const int d = mesh->ndims;
  const char* src = static_cast<const char*>(fv.ptr);
  int64_t lc[4] = {0, 0, 0, 0};
  for (uint64_t v = 0; v < nvert; ++v) {
    int64_t idx = 0;
    for (int j = 0; j < d; ++j) idx += lc[j] * fv.strides[j];
    std::memcpy(dst + v * esz, src + idx * (int64_t)esz, esz);
    for (int j = 0; j < d; ++j) {
      if (++lc[j] < (int64_t)mesh->local_dims[j]) break;
      lc[j] = 0;
    }
  }
}

// Element strides of a contiguous x-fastest copy of the local grid.
inline void packed_strides(const ftk2h_mesh_desc* mesh, int64_t* strides) {
  int64_t s = 1;
  for (int j = 0; j < 4; ++j) {
    strides[j] = j < mesh->ndims ? s : 0;
    if (j < mesh->ndims) s *= (int64_t)mesh->local_dims[j];
  }
}

// Per-corner offset tables, the packed-layout flag and the emit capacities, once
// the component and scalar views of P are bound.
inline void finish_params(KParams& P, const ftk2h_mesh_desc* mesh, int ncomp) {
  const int d = mesh->ndims;
  int64_t lst[4] = {0, 0, 0, 0}, ls = 1;
  for (int j = 0; j < d; ++j) {
    lst[j] = ls;
    ls *= (int64_t)mesh->local_dims[j];
  }
  P.packed = 1;
  for (int c = 0; c < ncomp; ++c)
    for (int j = 0; j < d; ++j)
      if (P.cst[c][j] != lst[j]) P.packed = 0;
  for (int s = 0; s < 16; ++s) {
    int64_t lo = 0, so = 0;
    uint64_t go = 0;
    int64_t co[3] = {0, 0, 0};
    for (int j = 0; j < d; ++j) {
      if (!((s >> j) & 1)) continue;
      lo += lst[j];
      go += (uint64_t)P.gstride[j];
      so += P.sst[j];
      for (int c = 0; c < 3; ++c) co[c] += P.cst[c][j];
    }
    P.loff[s] = lo;
    P.goff[s] = go;
    P.soff[s] = so;
    for (int c = 0; c < 3; ++c) P.coff[c][s] = co[c];
  }
  // Workgroup staging capacities of the emit kernel.  FTK2H_EMIT_CAPS="nq,words"
  // lowers them (tests use it to drive the overflow paths).
  P.nq_cap = EMIT_NQ;
  P.stage_cap = EMIT_STAGE_WORDS;
  if (const char* caps = std::getenv("FTK2H_EMIT_CAPS")) {
    unsigned a = 0, b = 0;
    if (std::sscanf(caps, "%u,%u", &a, &b) == 2) {
      P.nq_cap = a < (unsigned)EMIT_NQ ? a : EMIT_NQ;
      P.stage_cap = b < (unsigned)EMIT_STAGE_WORDS ? b : EMIT_STAGE_WORDS;
    }
  }
}

// Sort + unique the cells of R->conn[0] (bucket by smallest index, sort inside
// buckets, compact).  Replaces R->conn[0] and R->conn_cells[0].
template <int NV>
int sort_unique_cells(ftk2h_result* R, uint64_t nn, uint64_t nc, hipStream_t st) {
  uint32_t *cnt = nullptr, *sums = nullptr, *tmp = nullptr, *keep = nullptr, *pos = nullptr, *out = nullptr;
  const uint64_t tiles = (std::max(nn, nc) + 1 + SCAN_TILE - 1) / SCAN_TILE + 2;
  auto cleanup = [&]() {
    for (uint32_t* p : {cnt, sums, tmp, keep, pos})
      if (p) (void)hipFree(p);
  };
  auto check = [&](hipError_t e, const char* what) -> int {
    if (e == hipSuccess) return FTK2H_OK;
    cleanup();
    return fail(e == hipErrorOutOfMemory ? FTK2H_E_OOM : FTK2H_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  int rc;
  if ((rc = check(hipMalloc(&cnt, (nn + 1) * 4), "hipMalloc"))) return rc;
  if ((rc = check(hipMalloc(&sums, tiles * 4), "hipMalloc"))) return rc;
  if ((rc = check(hipMalloc(&tmp, nc * NV * 4), "hipMalloc"))) return rc;
  if ((rc = check(hipMalloc(&keep, (nc + 1) * 4), "hipMalloc"))) return rc;
  const unsigned gn = (unsigned)((nn + BLOCK) / BLOCK), gc = (unsigned)((nc + BLOCK - 1) / BLOCK);
  k_fill_u32<<<gn, BLOCK, 0, st>>>(cnt, 0u, nn + 1);
  k_cell_count<NV><<<gc, BLOCK, 0, st>>>(R->conn[0], nc, cnt);
  if ((rc = check(scan_u32(cnt, cnt, nn, sums, st), "scan"))) return rc;  // cnt -> bucket offsets
  // scatter with a cursor copy of the offsets (keep holds the cursor temporarily)
  uint32_t* cursor = keep;
  if (nc + 1 < nn + 1) {
    (void)hipFree(keep);
    keep = nullptr;
    if ((rc = check(hipMalloc(&keep, (nn + 1) * 4), "hipMalloc"))) return rc;
    cursor = keep;
  }
  if ((rc = check(hipMemcpyAsync(cursor, cnt, (nn + 1) * 4, hipMemcpyDeviceToDevice, st), "copy"))) return rc;
  k_cell_scatter<NV><<<gc, BLOCK, 0, st>>>(R->conn[0], nc, cursor, tmp);
  k_fill_u32<<<(unsigned)((nc + BLOCK) / BLOCK), BLOCK, 0, st>>>(keep, 0u, nc + 1);
  k_cell_bucket_sort<NV><<<gn, BLOCK, 0, st>>>(tmp, cnt, nn, keep);
  // positions of kept cells
  if ((rc = check(hipMalloc(&pos, (nc + 1) * 4), "hipMalloc"))) return rc;
  if ((rc = check(scan_u32(keep, pos, nc, sums, st), "scan"))) return rc;
  uint32_t nu = 0;
  if ((rc = check(hipMemcpyAsync(&nu, pos + nc, 4, hipMemcpyDeviceToHost, st), "copy"))) return rc;
  if ((rc = check(hipStreamSynchronize(st), "sync"))) return rc;
  if ((rc = check(hipMalloc(&out, (uint64_t)std::max<uint32_t>(nu, 1) * NV * 4), "hipMalloc"))) return rc;
  k_cell_compact<NV><<<gc, BLOCK, 0, st>>>(tmp, keep, pos, nc, out);
  if ((rc = check(hipStreamSynchronize(st), "sync"))) return rc;
  R->release(R->conn[0]);
  R->conn[0] = out;
  R->conn_cells[0] = nu;
  cleanup();
  return FTK2H_OK;
}

template <int KIND, int D, int M, class T>
int run_pipeline(RunArgs& A) {
  constexpr int NV = CellKind<KIND, D, M>::verts;
  KParams& P = A.P;
  ftk2h_ctx* ctx = A.ctx;
  hipStream_t st = A.st;
  const int64_t nb = A.nb;
  auto rec = [&](int i) { return hipEventRecord(ctx->ev[i], st); };
  uint32_t* scan_sums = A.fixed ? A.fixed->scan_sums : static_cast<uint32_t*>(ctx->scan_sums.p);
  // events around each kernel (phases: classify, stage 2, node scan, offsets, cell scan
  // + sync, emit, link, compress)
  HIP_TRY(rec(0));
  if (nb > 0) k_classify<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(1));
  if (nb > 0) k_stage2<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(2));
  HIP_TRY(scan_u32(P.bsum, P.bsum, (uint64_t)nb, scan_sums, st));
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(3));
  if (nb > 0) k_offsets<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(4));
  HIP_TRY(scan_u32(P.ebsum, P.ebsum, (uint64_t)nb, scan_sums, st));
  HIP_TRY(hipGetLastError());
  uint32_t totals[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(&totals[0], P.bsum + nb, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&totals[1], P.ebsum + nb, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const uint64_t nn = totals[0], nc = totals[1];
  if (A.node_cap && nn > A.node_cap)
    return fail(FTK2H_E_OVERFLOW, "node capacity " + std::to_string(A.node_cap) + " < required " +
                                      std::to_string(nn));
  if (A.cell_cap && nc > A.cell_cap)
    return fail(FTK2H_E_OVERFLOW, "cell capacity " + std::to_string(A.cell_cap) + " < required " +
                                      std::to_string(nc));
  ftk2h_result* R = A.res;
  R->n_nodes = nn;
  R->ntypes = Chains<D, M>::N;
  uint32_t* parent = nullptr;
  if (const FixedOut* F = A.fixed) {
    if (nn > F->node_cap || (NV > 0 && nc > F->cell_cap)) {
      R->conn_cells[0] = nc;  // required sizes for the caller
      return fail(FTK2H_E_OVERFLOW, "output capacity: " + std::to_string(nn) + " nodes and " +
                                        std::to_string(nc) + " cells required");
    }
    const uint64_t pbytes = (nn * 4 + 255) & ~(uint64_t)255;
    const uint64_t xbytes = NV > 0 ? nc * (uint64_t)(NV > 1 ? NV - 1 : 1) * 8 : 0;
    if (pbytes + xbytes > F->tail_bytes)
      return fail(FTK2H_E_OVERFLOW, "scratch capacity: " + std::to_string(nn) + " nodes and " + std::to_string(nc) +
                                        " cells need " + std::to_string(pbytes + xbytes) + " bytes, " +
                                        std::to_string(F->tail_bytes) + " left in the fixed scratch array");
    if (P.scal && !F->scalar) return fail(FTK2H_E_INVALID, "a scalar field needs a scalar output array");
    R->keys = F->keys;
    R->lambda = F->lambda;
    R->track = F->track;
    R->scalar = P.scal ? F->scalar : nullptr;
    parent = reinterpret_cast<uint32_t*>(F->tail);
    if (NV > 0 && nc) {
      P.xedges = reinterpret_cast<uint32_t*>(F->tail + pbytes);
      R->n_conn = 1;
      R->conn_dim[0] = NV - 1;
      R->conn_cells[0] = nc;
      R->conn[0] = F->cells;
    }
  } else {
  R->pool = ctx->pool;
  DevPool& pool = *ctx->pool;
  if (nn) {
    HIP_TRY(pool.get(reinterpret_cast<void**>(&R->keys), nn * 8));
    HIP_TRY(pool.get(&R->lambda, nn * (M + 1) * sizeof(T)));
    HIP_TRY(pool.get(reinterpret_cast<void**>(&R->track), nn * 4));
    if (P.scal) HIP_TRY(pool.get(reinterpret_cast<void**>(&R->scalar), nn * 4));
    if (int rc = ctx->parent.ensure(nn * 4)) return rc;
    parent = static_cast<uint32_t*>(ctx->parent.p);
  }
  if (NV > 0 && nc) {
    int rc = ctx->xedges.ensure(nc * (uint64_t)(NV - 1) * 8);
    if (rc) return rc;
    P.xedges = static_cast<uint32_t*>(ctx->xedges.p);
    R->n_conn = 1;
    R->conn_dim[0] = NV - 1;
    R->conn_cells[0] = nc;
    HIP_TRY(pool.get(reinterpret_cast<void**>(&R->conn[0]), nc * NV * 4));
  }
  }
  P.keys = R->keys;
  P.lambda = R->lambda;
  P.scalar_out = R->scalar;
  P.cells = R->conn[0];
  P.parent = parent;
  P.track = R->track;
  P.n_nodes = nn;
  P.n_cells = nc;
  HIP_TRY(rec(5));
  if (nb > 0 && nn) k_emit<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(6));
  if constexpr (NV > 0) {
    if (nc) {
      // at most nc * (NV - 1) pairs; threads beyond the device-side count exit
      const uint64_t g = (nc * (NV - 1) + BLOCK - 1) / BLOCK;
      k_link_pairs<NV != 2><<<dim3((unsigned)std::min<uint64_t>(g, LINK_GRID)), dim3(BLOCK), 0, st>>>(parent, P.xedges, P.xcount);
      HIP_TRY(hipGetLastError());
    }
  }
  HIP_TRY(rec(7));
  if (nn) {
    launch_compress(parent, R->track, nn, st);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(rec(8));
  if constexpr (CellKind<KIND, D, M>::value == 2) {
    // star-fan triangles: global sort + unique (FeatureComplex::sort)
    if (nc) {
      int rc = sort_unique_cells<NV>(R, nn, nc, st);
      if (rc) return rc;
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < NPHASE; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]) == hipSuccess) R->phase_ms[i] = ms;
  }
  return FTK2H_OK;
}

}  // namespace ftk2h_internal
