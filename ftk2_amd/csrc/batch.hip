// C ABI of batched extraction (ftk2h_batch_*, include/ftk2h.h): job table set-up,
// graph capture/replay and the schedule autotuner.  Kernels and the device side of
// a run are in batch.hpp (instantiated per predicate in inst_batch_*.hip).
#include "batch.hpp"

#include <chrono>

using namespace ftk2h;
using namespace ftk2h_internal;

namespace ftk2h_internal {

#define FTK2H_BATCH_EXTERN(KIND, D, M)                                                           \
  extern template int run_batch<KIND, D, M, float>(BatchArgs&, std::vector<ftk2h_result*>&); \
  extern template int run_batch<KIND, D, M, double>(BatchArgs&, std::vector<ftk2h_result*>&);
FTK2H_BATCH_EXTERN(KIND_CP, 4, 3)
FTK2H_BATCH_EXTERN(KIND_CP, 3, 2)
FTK2H_BATCH_EXTERN(KIND_CP, 2, 1)
FTK2H_BATCH_EXTERN(KIND_CONTOUR, 4, 1)
FTK2H_BATCH_EXTERN(KIND_CONTOUR, 3, 1)
FTK2H_BATCH_EXTERN(KIND_CONTOUR, 2, 1)
FTK2H_BATCH_EXTERN(KIND_FIBER, 3, 2)
#undef FTK2H_BATCH_EXTERN

using BatchFn = int (*)(BatchArgs&, std::vector<ftk2h_result*>&);

namespace {
struct BatchEntry {
  int kind, d, m, dtype;
  BatchFn fn;
};
template <int KIND, int D, int M>
constexpr BatchEntry bentry_f() {
  return {KIND, D, M, FTK2H_F32, &run_batch<KIND, D, M, float>};
}
template <int KIND, int D, int M>
constexpr BatchEntry bentry_d() {
  return {KIND, D, M, FTK2H_F64, &run_batch<KIND, D, M, double>};
}
const BatchEntry kBatchEntries[] = {
    bentry_d<KIND_CP, 4, 3>(),      bentry_f<KIND_CP, 4, 3>(),      bentry_d<KIND_CP, 3, 2>(),
    bentry_f<KIND_CP, 3, 2>(),      bentry_d<KIND_CP, 2, 1>(),      bentry_f<KIND_CP, 2, 1>(),
    bentry_d<KIND_CONTOUR, 4, 1>(), bentry_f<KIND_CONTOUR, 4, 1>(), bentry_d<KIND_CONTOUR, 3, 1>(),
    bentry_f<KIND_CONTOUR, 3, 1>(), bentry_d<KIND_CONTOUR, 2, 1>(), bentry_f<KIND_CONTOUR, 2, 1>(),
    bentry_d<KIND_FIBER, 3, 2>(),   bentry_f<KIND_FIBER, 3, 2>(),
};
BatchFn find_batch_entry(int kind, int d, int m, int dtype) {
  for (const auto& e : kBatchEntries)
    if (e.kind == kind && e.d == d && e.m == m && e.dtype == dtype) return e.fn;
  return nullptr;
}
}  // namespace

int launch_graph_or_direct(BatchState* S, hipGraphExec_t* g, std::vector<uint64_t>* sig,
                           const std::vector<uint64_t>& want, const std::function<int()>& body) {
  if (!S->cfg.use_graph) return body();
  if (*g && *sig == want) {
    HIP_TRY(hipGraphLaunch(*g, S->st));
    return FTK2H_OK;
  }
  if (*g) {
    (void)hipGraphExecDestroy(*g);
    *g = nullptr;
    sig->clear();
  }
  HIP_TRY(hipStreamBeginCapture(S->st, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(S->st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(FTK2H_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(FTK2H_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  *g = exec;
  *sig = want;
  HIP_TRY(hipGraphLaunch(exec, S->st));
  return FTK2H_OK;
}

}  // namespace ftk2h_internal

struct ftk2h_batch {
  BatchState S;
  BatchFn fn = nullptr;
  std::mutex lock;
};

namespace {

int run_locked(ftk2h_batch* b, int njobs, const ftk2h_job* jobs, ftk2h_result** results, void* hip_stream) {
  BatchState& S = b->S;
  const auto t_host0 = std::chrono::steady_clock::now();
  const int d = S.d;
  for (int j = 0; j < njobs; ++j) results[j] = nullptr;
  HIP_TRY(hipSetDevice(S.ctx->device));
  hipStream_t own = S.st;
  if (hip_stream) S.st = static_cast<hipStream_t>(hip_stream);
  struct Restore {
    BatchState& S;
    hipStream_t own;
    ~Restore() { S.st = own; }
  } restore{S, own};

  // ---- validate, make the result objects, lay out the active jobs
  std::vector<std::unique_ptr<ftk2h_result>> res(njobs);
  std::vector<int> active;
  std::vector<uint64_t> nverts(njobs, 0);
  for (int j = 0; j < njobs; ++j) {
    const ftk2h_job& J = jobs[j];
    if (!J.comps) return fail(FTK2H_E_INVALID, "job " + std::to_string(j) + ": null field views");
    for (int c = 0; c < S.ncomp; ++c)
      if (!J.comps[c].ptr) return fail(FTK2H_E_INVALID, "job " + std::to_string(j) + ": null field pointer");
    bool has_cells = false;
    if (int rc = check_mesh(&J.mesh, d, &nverts[j], &has_cells))
      return fail(rc, "job " + std::to_string(j) + ": " + error_slot());
    if (J.mesh.time_id_stride)
      return fail(FTK2H_E_UNSUPPORTED, "job " + std::to_string(j) + ": batches take regular meshes only");
    auto R = std::make_unique<ftk2h_result>();
    R->device = S.ctx->device;
    R->kind = S.pred.kind;
    R->d = d;
    R->m = S.m;
    R->dtype = S.pred.dtype;
    for (int k = 0; k < d; ++k) R->gdims[k] = J.mesh.global_dims[k];
    R->has_scalar = J.scalar && J.scalar->ptr && S.pred.kind == FTK2H_PRED_CRITICAL_POINT;
    if (S.pred.kind == FTK2H_PRED_CONTOUR) R->scalar_value = (float)S.pred.thresholds[0];
    res[j] = std::move(R);
    if (has_cells) active.push_back(j);
  }
  if (active.empty()) {
    for (int j = 0; j < njobs; ++j) results[j] = res[j].release();
    return FTK2H_OK;
  }
  const uint32_t na = (uint32_t)active.size();
  std::vector<uint64_t> tile_base(na + 1, 0), vert_base(na + 1, 0);
  uint64_t staging_bytes = 0;
  bool any_scalar = false;
  for (uint32_t a = 0; a < na; ++a) {
    const int j = active[a];
    const uint64_t nt = (nverts[j] + BLOCK - 1) / BLOCK;
    tile_base[a + 1] = tile_base[a] + nt;
    vert_base[a + 1] = vert_base[a] + nverts[j];
    for (int c = 0; c < S.ncomp; ++c)
      if (!jobs[j].comps[c].on_device) staging_bytes += (nverts[j] * S.esz + 255) & ~(uint64_t)255;
    if (res[j]->has_scalar) {
      any_scalar = true;
      if (!jobs[j].scalar->on_device) staging_bytes += (nverts[j] * S.esz + 255) & ~(uint64_t)255;
    }
  }
  const uint64_t ntiles = tile_base[na], totv = vert_base[na];
  if (ntiles >= 0xFFFFFFFFull) return fail(FTK2H_E_OVERFLOW, "batch: too many tiles");

  int rc;
  if ((rc = S.jobs.ensure((size_t)na * sizeof(KParams)))) return rc;
  if ((rc = S.tile_base.ensure((na + 1) * 8))) return rc;
  if ((rc = S.tile_job.ensure(ntiles * 4))) return rc;
  if ((rc = S.mask.ensure(totv * 8))) return rc;
  if ((rc = S.pend.ensure(totv * 8))) return rc;
  if ((rc = S.degm.ensure(totv * 8))) return rc;
  if ((rc = S.ccount.ensure(totv * 2))) return rc;
  if ((rc = S.node_off.ensure(totv * 4))) return rc;
  if ((rc = S.bsum.ensure((ntiles + 1) * 4))) return rc;
  if ((rc = S.ebsum.ensure((ntiles + 1) * 4))) return rc;
  if ((rc = S.sums.ensure((ntiles / SCAN_TILE + 2) * 4))) return rc;
  if ((rc = S.counters.ensure(256))) return rc;
  if ((rc = S.bases.ensure((size_t)(na + 1) * 2 * 4))) return rc;
  if (staging_bytes && (rc = S.staging.ensure(staging_bytes))) return rc;
  if (S.host_bases_cap < (size_t)(na + 1) * 2) {
    if (S.host_bases) (void)hipHostFree(S.host_bases);
    S.host_bases = nullptr;
    S.host_bases_cap = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&S.host_bases), (size_t)(na + 1) * 2 * 4, hipHostMallocDefault));
    S.host_bases_cap = (size_t)(na + 1) * 2;
  }

  // ---- job table
  std::vector<KParams> table(na);
  uint64_t stage_off = 0;
  std::vector<char> pack;
  auto bind = [&](const ftk2h_field_view& fv, const ftk2h_mesh_desc* mesh, uint64_t nvert, const void** ptr,
                  int64_t* strides) -> int {
    if (fv.on_device) {
      *ptr = fv.ptr;
      for (int k = 0; k < 4; ++k) strides[k] = k < d ? fv.strides[k] : 0;
      return FTK2H_OK;
    }
    pack.resize(nvert * S.esz);
    pack_host_view(fv, mesh, S.esz, nvert, pack.data());
    char* dst = static_cast<char*>(S.staging.p) + stage_off;
    HIP_TRY(hipMemcpyAsync(dst, pack.data(), pack.size(), hipMemcpyHostToDevice, S.st));
    HIP_TRY(hipStreamSynchronize(S.st));  // pack is reused
    stage_off += (pack.size() + 255) & ~(uint64_t)255;
    *ptr = dst;
    packed_strides(mesh, strides);
    return FTK2H_OK;
  };
  for (uint32_t a = 0; a < na; ++a) {
    const int j = active[a];
    const ftk2h_job& J = jobs[j];
    KParams& P = table[a];
    init_params(P, &J.mesh, &S.pred, nverts[j]);
    for (int c = 0; c < S.ncomp; ++c)
      if ((rc = bind(J.comps[c], &J.mesh, nverts[j], &P.comp[c], P.cst[c]))) return rc;
    if (res[j]->has_scalar && (rc = bind(*J.scalar, &J.mesh, nverts[j], &P.scal, P.sst))) return rc;
    finish_params(P, &J.mesh, S.ncomp);
    P.mask = static_cast<uint64_t*>(S.mask.p) + vert_base[a];
    P.pend = static_cast<uint64_t*>(S.pend.p) + vert_base[a];
    P.degm = static_cast<uint64_t*>(S.degm.p) + vert_base[a];
    P.ccount = static_cast<uint16_t*>(S.ccount.p) + vert_base[a];
    P.node_off = static_cast<uint32_t*>(S.node_off.p) + vert_base[a];
    P.bsum = static_cast<uint32_t*>(S.bsum.p) + tile_base[a];
    P.ebsum = static_cast<uint32_t*>(S.ebsum.p) + tile_base[a];
    P.xcount = static_cast<uint32_t*>(S.counters.p) + 3;
  }
  HIP_TRY(hipMemcpyAsync(S.jobs.p, table.data(), (size_t)na * sizeof(KParams), hipMemcpyHostToDevice, S.st));
  HIP_TRY(hipMemcpyAsync(S.tile_base.p, tile_base.data(), (na + 1) * 8, hipMemcpyHostToDevice, S.st));
  kb_tile_jobs<<<dim3((unsigned)((ntiles + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, S.st>>>(
      static_cast<const uint64_t*>(S.tile_base.p), na, ntiles, static_cast<uint32_t*>(S.tile_job.p));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(S.ev[0], S.st));

  std::vector<ftk2h_result*> outs(na);
  for (uint32_t a = 0; a < na; ++a) outs[a] = res[active[a]].get();
  BatchArgs A{&S, (int)na, ntiles, totv, any_scalar};
  if ((rc = b->fn(A, outs))) return rc;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, S.ev[0], S.ev[1]) == hipSuccess) S.last_ms[0] = ms;
  if (hipEventElapsedTime(&ms, S.ev[1], S.ev[2]) == hipSuccess) S.last_ms[1] = ms;
  S.last_ms[2] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
  for (int j = 0; j < njobs; ++j) results[j] = res[j].release();
  return FTK2H_OK;
}

}  // namespace

extern "C" {

int ftk2h_batch_create(ftk2h_ctx* ctx, const ftk2h_predicate_desc* pred, int ndims, ftk2h_batch** out) {
  if (!ctx || !pred || !out) return fail(FTK2H_E_INVALID, "null argument");
  *out = nullptr;
  const int need = components_needed(pred);
  if (need < 0) return FTK2H_E_INVALID;
  if (pred->dtype != FTK2H_F32 && pred->dtype != FTK2H_F64) return fail(FTK2H_E_INVALID, "bad dtype");
  BatchFn fn = find_batch_entry(pred->kind, ndims, pred->codimension, pred->dtype);
  if (!fn)
    return fail(FTK2H_E_UNSUPPORTED, "batch: predicate kind " + std::to_string(pred->kind) + " with d=" +
                                         std::to_string(ndims) + " m=" + std::to_string(pred->codimension) +
                                         " is not batched");
  auto b = std::make_unique<ftk2h_batch>();
  b->fn = fn;
  BatchState& S = b->S;
  S.ctx = ctx;
  S.pred = *pred;
  S.d = ndims;
  S.m = pred->codimension;
  S.ncomp = need;
  S.esz = pred->dtype == FTK2H_F64 ? 8 : 4;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceGetAttribute(&S.n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  HIP_TRY(hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking));
  for (auto& e : S.ev) HIP_TRY(hipEventCreate(&e));
  *out = b.release();
  return FTK2H_OK;
}

int ftk2h_batch_set_config(ftk2h_batch* b, const ftk2h_batch_config* cfg) {
  if (!b || !cfg) return fail(FTK2H_E_INVALID, "null argument");
  if (cfg->grid_per_cu < 0 || cfg->grid_per_cu > 64 || cfg->chunk < 0 || cfg->chunk > 4096)
    return fail(FTK2H_E_INVALID, "batch config out of range");
  std::lock_guard<std::mutex> g(b->lock);
  b->S.cfg.grid_per_cu = cfg->grid_per_cu;
  b->S.cfg.chunk = cfg->chunk;
  b->S.cfg.use_graph = cfg->use_graph ? 1 : 0;
  return FTK2H_OK;
}

int ftk2h_batch_get_config(const ftk2h_batch* b, ftk2h_batch_config* cfg) {
  if (!b || !cfg) return fail(FTK2H_E_INVALID, "null argument");
  cfg->grid_per_cu = b->S.cfg.grid_per_cu;
  cfg->chunk = b->S.cfg.chunk;
  cfg->use_graph = b->S.cfg.use_graph;
  return FTK2H_OK;
}

int ftk2h_batch_run(ftk2h_batch* b, int njobs, const ftk2h_job* jobs, ftk2h_result** results, void* hip_stream) {
  if (!b || njobs < 0 || (njobs && (!jobs || !results))) return fail(FTK2H_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(b->lock);
  return run_locked(b, njobs, jobs, results, hip_stream);
}

int ftk2h_batch_autotune(ftk2h_batch* b, int njobs, const ftk2h_job* jobs, int reps, ftk2h_batch_config* best,
                         float* ms) {
  if (!b || njobs <= 0 || !jobs) return fail(FTK2H_E_INVALID, "null argument");
  if (reps < 1) reps = 1;
  std::lock_guard<std::mutex> g(b->lock);
  static const int kCand[][2] = {{0, 0}, {1, 0}, {2, 0}, {4, 0}, {8, 0}, {2, 1}, {2, 4},
                                 {4, 1}, {4, 4}, {4, 16}, {8, 4}, {8, 16}, {16, 4}};
  const BatchConfig keep = b->S.cfg;
  BatchConfig win = keep;
  float win_ms = 0.f;
  std::vector<ftk2h_result*> out(njobs, nullptr);
  for (const auto& c : kCand) {
    b->S.cfg.grid_per_cu = c[0];
    b->S.cfg.chunk = c[1];
    b->S.cfg.use_graph = keep.use_graph;
    float cand = 0.f;
    for (int r = 0; r <= reps; ++r) {  // run 0 captures the graphs of this shape; untimed
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = run_locked(b, njobs, jobs, out.data(), nullptr);
      const float dt = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
      for (auto*& res : out) {
        delete res;
        res = nullptr;
      }
      if (rc) {
        b->S.cfg = keep;
        return rc;
      }
      if (r == 1 || (r > 1 && dt < cand)) cand = dt;
    }
    if (win_ms == 0.f || cand < win_ms) {
      win_ms = cand;
      win = b->S.cfg;
    }
  }
  b->S.cfg = win;
  if (best) {
    best->grid_per_cu = win.grid_per_cu;
    best->chunk = win.chunk;
    best->use_graph = win.use_graph;
  }
  if (ms) *ms = win_ms;
  return FTK2H_OK;
}

int ftk2h_batch_last_ms(const ftk2h_batch* b, float* ms, int n) {
  if (!b || !ms) return fail(FTK2H_E_INVALID, "null argument");
  for (int i = 0; i < n && i < 3; ++i) ms[i] = b->S.last_ms[i];
  return FTK2H_OK;
}

void ftk2h_batch_destroy(ftk2h_batch* b) {
  if (!b) return;
  BatchState& S = b->S;
  (void)hipSetDevice(S.ctx->device);
  if (S.st) (void)hipStreamSynchronize(S.st);
  if (S.g1) (void)hipGraphExecDestroy(S.g1);
  if (S.g2) (void)hipGraphExecDestroy(S.g2);
  for (DevBuf* p : {&S.jobs, &S.tile_base, &S.tile_job, &S.mask, &S.pend, &S.degm, &S.ccount, &S.node_off, &S.bsum, &S.ebsum, &S.sums,
                    &S.counters, &S.bases, &S.xedges, &S.parent, &S.staging})
    p->release();
  if (S.host_bases) (void)hipHostFree(S.host_bases);
  S.arena.reset();  // results that still hold it keep it alive
  for (auto& e : S.ev)
    if (e) (void)hipEventDestroy(e);
  if (S.st) (void)hipStreamDestroy(S.st);
  delete b;
}

}  // extern "C"
