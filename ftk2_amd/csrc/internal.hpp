// Internal declarations shared by the host translation units of libftk2h:
// error reporting, device buffers, context/result objects and the templated
// pipeline driver (instantiated per predicate in inst_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ftk2h.h"
#include "kernels.hpp"
#include "sortcells.hpp"

#include <algorithm>

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

constexpr int NPHASE = 7;

}  // namespace ftk2h_internal

struct ftk2h_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  ftk2h_internal::DevBuf mask, node_off, bsum, ebsum, xedges, xcount, fields[4];
  hipEvent_t ev[ftk2h_internal::NPHASE + 1] = {};
  std::mutex lock;
};

struct ftk2h_result {
  int device = 0;
  int kind = 0, d = 0, m = 0, dtype = 0;
  uint64_t gdims[4] = {1, 1, 1, 1};
  uint64_t n_nodes = 0;
  int ntypes = 0;
  uint64_t* keys = nullptr;
  void* lambda = nullptr;
  uint32_t* track = nullptr;
  float* scalar = nullptr;
  int n_conn = 0;
  int conn_dim[3] = {0, 0, 0};
  uint64_t conn_cells[3] = {0, 0, 0};
  uint32_t* conn[3] = {nullptr, nullptr, nullptr};
  float phase_ms[ftk2h_internal::NPHASE] = {};
  ~ftk2h_result() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    if (keys) (void)hipFree(keys);
    if (lambda) (void)hipFree(lambda);
    if (track) (void)hipFree(track);
    if (scalar) (void)hipFree(scalar);
    for (auto* c : conn)
      if (c) (void)hipFree(c);
    (void)hipSetDevice(cur);
  }
};

namespace ftk2h_internal {

struct RunArgs {
  ftk2h_ctx* ctx;
  hipStream_t st;
  KParams P;
  int64_t nb;  // blocks over vertices
  uint64_t node_cap, cell_cap;
  bool sync;
  ftk2h_result* res;
};

using RunFn = int (*)(RunArgs&);

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
  (void)hipFree(R->conn[0]);
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
  HIP_TRY(rec(0));
  if (nb > 0) k_classify<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(1));
  k_scan_small<<<1, 1024, 0, st>>>(P.bsum, nb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(2));
  if (nb > 0) k_offsets<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(3));
  k_scan_small<<<1, 1024, 0, st>>>(P.ebsum, nb);
  HIP_TRY(hipGetLastError());
  uint32_t totals[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(&totals[0], P.bsum + nb, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&totals[1], P.ebsum + nb, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(rec(4));
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
  if (nn) {
    HIP_TRY(hipMalloc(&R->keys, nn * 8));
    HIP_TRY(hipMalloc(&R->lambda, nn * (M + 1) * sizeof(T)));
    HIP_TRY(hipMalloc(&R->track, nn * 4));
    HIP_TRY(hipMalloc(&parent, nn * 4));
    if (P.scal) HIP_TRY(hipMalloc(&R->scalar, nn * 4));
  }
  if (NV > 0 && nc) {
    int rc = ctx->xedges.ensure(nc * (uint64_t)(NV - 1) * 8);
    if (rc) return rc;
    P.xedges = static_cast<uint32_t*>(ctx->xedges.p);
    R->n_conn = 1;
    R->conn_dim[0] = NV - 1;
    R->conn_cells[0] = nc;
    HIP_TRY(hipMalloc(&R->conn[0], nc * NV * 4));
  }
  P.keys = R->keys;
  P.lambda = R->lambda;
  P.scalar_out = R->scalar;
  P.cells = R->conn[0];
  P.parent = parent;
  P.track = R->track;
  P.n_nodes = nn;
  P.n_cells = nc;
  if (nb > 0 && nn) k_emit<KIND, D, M, T><<<dim3((unsigned)nb), dim3(BLOCK), 0, st>>>(P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(rec(5));
  if constexpr (NV > 0) {
    if (nc) {
      // at most nc * (NV - 1) pairs; threads beyond the device-side count exit
      const uint64_t g = (nc * (NV - 1) + BLOCK - 1) / BLOCK;
      k_link_pairs<<<dim3((unsigned)g), dim3(BLOCK), 0, st>>>(parent, P.xedges, P.xcount);
      HIP_TRY(hipGetLastError());
    }
  }
  HIP_TRY(rec(6));
  if (nn) {
    const uint64_t g = (nn + BLOCK - 1) / BLOCK;
    k_compress<<<dim3((unsigned)g), dim3(BLOCK), 0, st>>>(parent, R->track, nn);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(rec(7));
  if constexpr (CellKind<KIND, D, M>::value == 2) {
    // star-fan triangles: global sort + unique (FeatureComplex::sort)
    if (nc) {
      int rc = sort_unique_cells<NV>(R, nn, nc, st);
      if (rc) return rc;
    }
  }
  // parent is scratch; free after the stream drains
  HIP_TRY(hipStreamSynchronize(st));
  if (parent) HIP_TRY(hipFree(parent));
  for (int i = 0; i < NPHASE; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]) == hipSuccess) R->phase_ms[i] = ms;
  }
  return FTK2H_OK;
}

}  // namespace ftk2h_internal
