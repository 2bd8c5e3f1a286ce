// Batched extraction: many RegularSimplicialMesh jobs of one predicate template in
// one pass over the device.  Every job keeps its own KParams (in a device job
// table); the jobs' 256-vertex tiles are numbered consecutively, so the per-tile
// node and cell counts of all jobs form one array and one device-wide scan gives
// every node and cell its position in shared output arrays.  Node indices are
// batch-global while the kernels run (emit, union-find) and are made job-local
// (FeatureComplex indices of that job) by the last kernels.
//
// Scheduling (the autotuned part): tiles are handed to workgroups either one per
// workgroup (grid = tiles), by a static grid stride over a persistent grid, or
// dynamically in chunks of consecutive tiles taken from an atomic counter.
#pragma once
#include "internal.hpp"

namespace ftk2h {

struct BatchDev {
  const KParams* jobs;        // njobs entries
  const uint32_t* tile_job;   // job of every tile (ntiles)
  const uint64_t* tile_base;  // first tile of every job (njobs + 1)
  uint64_t ntiles;
  uint32_t* counters;         // dynamic schedule: one counter per kernel (0 classify, 1 offsets,
                              // 2 emit, 4 stage 2); [3] counts cross-workgroup edges
  uint32_t chunk;             // tiles per grab; 0 = static grid stride
};

// Runs f(P, local tile) for every tile given to this workgroup.  The job table
// comes in as its own restrict-qualified kernel argument: nothing the kernel
// stores can alias it, so the compiler reads a job's parameters with scalar loads
// (as it reads kernel arguments) instead of per-lane vector loads.
template <class F>
__device__ __forceinline__ void for_tiles(const BatchDev& B, const KParams* __restrict__ jobs, int phase, F&& f) {
  // one call site of f, so that the tile body is inlined into the kernel
  __shared__ uint32_t grab;
  const bool dyn = B.chunk != 0;
  uint64_t t = dyn ? 0 : blockIdx.x, end = 0;
  while (true) {
    if (dyn && t >= end) {  // take the next chunk of consecutive tiles
      __syncthreads();
      if (threadIdx.x == 0) grab = atomicAdd(B.counters + phase, B.chunk);
      __syncthreads();
      t = __builtin_amdgcn_readfirstlane(grab);
      end = t + B.chunk;
    }
    if (t >= B.ntiles) break;
    __syncthreads();  // LDS of the previous tile is no longer read
    const uint32_t j = B.tile_job[t];
    f(jobs[j], (int64_t)(t - B.tile_base[j]));
    t += dyn ? 1 : gridDim.x;
  }
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK, 3) void kb_classify(BatchDev B, const KParams* __restrict__ jobs) {
  for_tiles(B, jobs, 0, [](const KParams& P, int64_t blk) __attribute__((always_inline)) { classify_tile<KIND, D, M, T>(P, blk); });
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK, 4) void kb_stage2(BatchDev B, const KParams* __restrict__ jobs) {
  for_tiles(B, jobs, 4, [](const KParams& P, int64_t blk) __attribute__((always_inline)) { stage2_tile<KIND, D, M, T>(P, blk); });
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void kb_offsets(BatchDev B, const KParams* __restrict__ jobs) {
  for_tiles(B, jobs, 1, [](const KParams& P, int64_t blk) __attribute__((always_inline)) { offsets_tile<KIND, D, M, T>(P, blk); });
}

template <int KIND, int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void kb_emit(BatchDev B, const KParams* __restrict__ jobs) {
  for_tiles(B, jobs, 2, [](const KParams& P, int64_t blk) __attribute__((always_inline)) { emit_tile<KIND, D, M, T>(P, blk); });
}

// tile -> job map
static __global__ __launch_bounds__(BLOCK) void kb_tile_jobs(const uint64_t* tile_base, uint32_t njobs,
                                                            uint64_t ntiles, uint32_t* tile_job) {
  const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t lo = 0, hi = njobs;  // last job with tile_base <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (tile_base[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  tile_job[t] = lo;
}

// Zero the schedule counters (a kernel, so that the sequence can be captured).
static __global__ void kb_zero(uint32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Node and cell bases of every job (scanned tile counts at the job's first tile).
static __global__ __launch_bounds__(BLOCK) void kb_job_bases(const uint64_t* tile_base, uint32_t njobs,
                                                            const uint32_t* bsum, const uint32_t* ebsum,
                                                            uint32_t* node_base, uint32_t* cell_base) {
  const uint32_t j = blockIdx.x * BLOCK + threadIdx.x;
  if (j > njobs) return;
  node_base[j] = bsum[tile_base[j]];
  cell_base[j] = ebsum[tile_base[j]];
}

// Point every job at the shared output arrays.
static __global__ __launch_bounds__(BLOCK) void kb_bind_outputs(KParams* jobs, uint32_t njobs, uint64_t* keys,
                                                               void* lambda, float* scalar, uint32_t* cells,
                                                               uint32_t* xedges, uint32_t* parent,
                                                               uint32_t* track, uint64_t n_nodes,
                                                               uint64_t n_cells) {
  const uint32_t j = blockIdx.x * BLOCK + threadIdx.x;
  if (j >= njobs) return;
  KParams& P = jobs[j];
  P.keys = keys;
  P.lambda = lambda;
  P.scalar_out = P.scal ? scalar : nullptr;
  P.cells = cells;
  P.xedges = xedges;
  P.parent = parent;
  P.track = track;
  P.n_nodes = n_nodes;
  P.n_cells = n_cells;
}

__device__ __forceinline__ uint32_t job_of(const uint32_t* base, uint32_t njobs, uint32_t i) {
  uint32_t lo = 0, hi = njobs;  // last job with base <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (base[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// track = root (component minimum) made job-local.  Jobs are disjoint ranges of
// the node array and no cell crosses jobs, so every root lies in its node's job.
static __global__ __launch_bounds__(BLOCK) void kb_compress(const uint32_t* parent, uint32_t* track,
                                                           const uint32_t* node_base, uint32_t njobs,
                                                           uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  uint32_t x = (uint32_t)i;
  uint32_t p = parent[x];
  while (p != x) {
    x = p;
    p = parent[x];
  }
  track[i] = x - node_base[job_of(node_base, njobs, (uint32_t)i)];
}

template <int NV>
__global__ __launch_bounds__(BLOCK) void kb_localize_cells(uint32_t* cells, const uint32_t* cell_base,
                                                           const uint32_t* node_base, uint32_t njobs,
                                                           uint64_t nc) {
  const uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc) return;
  const uint32_t b = node_base[job_of(cell_base, njobs, (uint32_t)c)];
#pragma unroll
  for (int q = 0; q < NV; ++q) cells[c * NV + q] -= b;
}

}  // namespace ftk2h

namespace ftk2h_internal {

// Device memory shared by the results of one batch run.
struct Arena {
  int device = 0;
  void* base = nullptr;
  size_t bytes = 0;
  ~Arena() {
    if (!base) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(base);
    (void)hipSetDevice(cur);
  }
};

struct BatchConfig {
  int grid_per_cu = 0;  // 0: one workgroup per tile
  int chunk = 0;        // 0: static stride (persistent grid only)
  int use_graph = 1;
};

// Host-side state of a batch object (per predicate template); buffers grow only.
struct BatchState {
  ftk2h_ctx* ctx = nullptr;
  ftk2h_predicate_desc pred{};
  int d = 0, m = 0, ncomp = 0, nv = 0;
  size_t esz = 8;
  hipStream_t st = nullptr;
  int n_cu = 256;
  BatchConfig cfg;
  DevBuf jobs, tile_base, tile_job, mask, pend, degm, ccount, node_off, bsum, ebsum, sums, counters, bases, xedges, parent, staging;
  uint32_t* host_bases = nullptr;  // pinned: node_base[njobs+1], cell_base[njobs+1]
  size_t host_bases_cap = 0;
  std::shared_ptr<Arena> arena;
  hipEvent_t ev[5] = {};
  float last_ms[4] = {};  // classify..offsets, emit..end, whole, host
  // captured launch sequences, replayed while their arguments are unchanged
  hipGraphExec_t g1 = nullptr, g2 = nullptr;
  std::vector<uint64_t> sig1, sig2;
};

struct BatchArgs {
  BatchState* S;
  int njobs;          // jobs with work (others already answered)
  uint64_t ntiles;
  uint64_t total_vert;
  bool any_scalar;
};

int launch_graph_or_direct(BatchState* S, hipGraphExec_t* g, std::vector<uint64_t>* sig,
                           const std::vector<uint64_t>& want, const std::function<int()>& body);

template <int KIND, int D, int M, class T>
int run_batch(BatchArgs& A, std::vector<ftk2h_result*>& outs) {
  constexpr int NV = CellKind<KIND, D, M>::verts;
  static_assert(CellKind<KIND, D, M>::value != 2, "star-fan cells are not batched");
  BatchState* S = A.S;
  hipStream_t st = S->st;
  const uint32_t nj = (uint32_t)A.njobs;
  const uint64_t ntiles = A.ntiles;
  BatchDev B;
  B.jobs = static_cast<const KParams*>(S->jobs.p);
  B.tile_job = static_cast<const uint32_t*>(S->tile_job.p);
  B.tile_base = static_cast<const uint64_t*>(S->tile_base.p);
  B.ntiles = ntiles;
  B.counters = static_cast<uint32_t*>(S->counters.p);
  B.chunk = S->cfg.grid_per_cu > 0 ? (uint32_t)S->cfg.chunk : 0u;
  const uint64_t grid_tiles = ntiles;
  const uint64_t grid_p = S->cfg.grid_per_cu > 0 ? std::min<uint64_t>(ntiles, (uint64_t)S->cfg.grid_per_cu * S->n_cu)
                                                 : grid_tiles;
  if (S->cfg.grid_per_cu <= 0) B.chunk = 0;  // one tile per workgroup
  uint32_t* bsum = static_cast<uint32_t*>(S->bsum.p);
  uint32_t* ebsum = static_cast<uint32_t*>(S->ebsum.p);
  uint32_t* sums = static_cast<uint32_t*>(S->sums.p);
  uint32_t* node_base = static_cast<uint32_t*>(S->bases.p);
  uint32_t* cell_base = node_base + (nj + 1);

  // ---- phase 1: classify, scan, offsets, scan, job bases -> host
  std::vector<uint64_t> sig{1, (uint64_t)(uintptr_t)S->jobs.p, (uint64_t)(uintptr_t)S->bsum.p,
                            (uint64_t)(uintptr_t)S->ebsum.p, (uint64_t)(uintptr_t)S->bases.p,
                            (uint64_t)(uintptr_t)S->host_bases, (uint64_t)(uintptr_t)S->tile_job.p,
                            (uint64_t)(uintptr_t)S->sums.p, ntiles, nj, grid_p, B.chunk};
  int rc = launch_graph_or_direct(S, &S->g1, &S->sig1, sig, [&]() -> int {
    kb_zero<<<1, 64, 0, st>>>(B.counters, 5);
    kb_classify<KIND, D, M, T><<<dim3((unsigned)grid_p), dim3(BLOCK), 0, st>>>(B, B.jobs);
    kb_stage2<KIND, D, M, T><<<dim3((unsigned)grid_p), dim3(BLOCK), 0, st>>>(B, B.jobs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(scan_u32(bsum, bsum, ntiles, sums, st));
    kb_offsets<KIND, D, M, T><<<dim3((unsigned)grid_p), dim3(BLOCK), 0, st>>>(B, B.jobs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(scan_u32(ebsum, ebsum, ntiles, sums, st));
    kb_job_bases<<<dim3((nj + BLOCK) / BLOCK), dim3(BLOCK), 0, st>>>(B.tile_base, nj, bsum, ebsum, node_base,
                                                                     cell_base);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(S->host_bases, node_base, (size_t)(nj + 1) * 2 * 4, hipMemcpyDeviceToHost, st));
    return FTK2H_OK;
  });
  if (rc) return rc;
  HIP_TRY(hipEventRecord(S->ev[1], st));
  HIP_TRY(hipStreamSynchronize(st));
  const uint32_t* hb = S->host_bases;
  const uint64_t nn = hb[nj], nc = hb[2 * (nj + 1) - 1];
  if (nn >= 0xFFFFFFFFull || nc >= 0xFFFFFFFFull) return fail(FTK2H_E_OVERFLOW, "batch: more than 2^32-1 nodes or cells");

  // ---- outputs: one arena for every job's arrays; scratch stays with the batch
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_keys = align(nn * 8), b_lam = align(nn * (M + 1) * sizeof(T)), b_track = align(nn * 4),
               b_scal = A.any_scalar ? align(nn * 4) : 0, b_cells = NV > 0 ? align(nc * NV * 4) : 0;
  const size_t need = b_keys + b_lam + b_track + b_scal + b_cells + 256;
  if (!S->arena || S->arena.use_count() > 1 || S->arena->bytes < need) {
    if (S->arena && S->arena.use_count() == 1) S->arena.reset();  // free before allocating anew
    auto a = std::make_shared<Arena>();
    a->device = S->ctx->device;
    const size_t want = need + need / 4;
    HIP_TRY(hipMalloc(&a->base, want));
    a->bytes = want;
    S->arena = a;
  }
  char* ab = static_cast<char*>(S->arena->base);
  uint64_t* keys = reinterpret_cast<uint64_t*>(ab);
  T* lambda = reinterpret_cast<T*>(ab + b_keys);
  uint32_t* track = reinterpret_cast<uint32_t*>(ab + b_keys + b_lam);
  float* scal = b_scal ? reinterpret_cast<float*>(ab + b_keys + b_lam + b_track) : nullptr;
  uint32_t* cells = b_cells ? reinterpret_cast<uint32_t*>(ab + b_keys + b_lam + b_track + b_scal) : nullptr;
  if ((rc = S->parent.ensure(std::max<uint64_t>(nn, 1) * 4))) return rc;
  if ((rc = S->xedges.ensure(std::max<uint64_t>(nc * (NV > 1 ? NV - 1 : 1), 1) * 8))) return rc;
  uint32_t* parent = static_cast<uint32_t*>(S->parent.p);
  uint32_t* xedges = static_cast<uint32_t*>(S->xedges.p);
  uint32_t* xcount = B.counters + 3;

  // ---- phase 2: bind outputs, emit, link, compress + localize
  std::vector<uint64_t> sig2{2, (uint64_t)(uintptr_t)S->jobs.p, (uint64_t)(uintptr_t)ab, nn, nc,
                             (uint64_t)(uintptr_t)parent, (uint64_t)(uintptr_t)xedges, ntiles, nj, grid_p,
                             B.chunk, (uint64_t)A.any_scalar};
  rc = launch_graph_or_direct(S, &S->g2, &S->sig2, sig2, [&]() -> int {
    kb_bind_outputs<<<dim3((nj + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st>>>(
        static_cast<KParams*>(S->jobs.p), nj, keys, lambda, scal, cells, xedges, parent, track, nn, nc);
    HIP_TRY(hipGetLastError());
    if (nn) {
      kb_emit<KIND, D, M, T><<<dim3((unsigned)grid_p), dim3(BLOCK), 0, st>>>(B, B.jobs);
      HIP_TRY(hipGetLastError());
      if constexpr (NV > 1) {
        if (nc) {
          const uint64_t g = (nc * (NV - 1) + BLOCK - 1) / BLOCK;
          k_link_pairs<NV != 2><<<dim3((unsigned)std::min<uint64_t>(g, LINK_GRID)), dim3(BLOCK), 0, st>>>(parent, xedges, xcount);
          HIP_TRY(hipGetLastError());
          kb_localize_cells<NV><<<dim3((unsigned)((nc + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st>>>(
              cells, cell_base, node_base, nj, nc);
          HIP_TRY(hipGetLastError());
        }
      }
      kb_compress<<<dim3((unsigned)((nn + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st>>>(parent, track, node_base,
                                                                                      nj, nn);
      HIP_TRY(hipGetLastError());
    }
    return FTK2H_OK;
  });
  if (rc) return rc;
  HIP_TRY(hipEventRecord(S->ev[2], st));
  HIP_TRY(hipStreamSynchronize(st));

  // ---- per-job results (views into the arena)
  for (uint32_t j = 0; j < nj; ++j) {
    ftk2h_result* R = outs[j];
    const uint64_t n0 = hb[j], n1 = hb[j + 1];
    const uint64_t c0 = hb[nj + 1 + j], c1 = hb[nj + 1 + j + 1];
    R->arena = S->arena;
    R->n_nodes = n1 - n0;
    R->ntypes = Chains<D, M>::N;
    if (R->n_nodes) {
      R->keys = keys + n0;
      R->lambda = lambda + n0 * (M + 1);
      R->track = track + n0;
      R->scalar = R->has_scalar && scal ? scal + n0 : nullptr;
    }
    if (NV > 0 && c1 > c0) {
      R->n_conn = 1;
      R->conn_dim[0] = NV - 1;
      R->conn_cells[0] = c1 - c0;
      R->conn[0] = cells + c0 * NV;
    }
  }
  return FTK2H_OK;
}

}  // namespace ftk2h_internal
