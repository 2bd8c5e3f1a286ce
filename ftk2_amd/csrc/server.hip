// Serving (ftk2h_server_*, include/ftk2h.h): streaming-tracking requests admitted
// against a device-memory budget and advanced by continuous batching.  Each
// request is an ftk2h_stream (execute_stream semantics: last timestep and open
// union-find frontier stay on the device).  A step takes every admitted stream
// that has a queued timestep, forms its next slab, and extracts the slabs of one
// predicate template in one batched pass (ftk2h_batch_run), except slabs large
// enough to fill the device and slabs of templates that are not batched (star-fan
// cells), which run one by one; streams join at the
// first step after admission and leave when closed and drained.  Because a slab's
// extraction is independent of the other jobs in its batch, each result equals
// the one the stream gets alone.
#include "stream.hpp"

#include <deque>
#include <map>
#include <thread>
#include <tuple>

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

// Slabs of at least this many vertices fill the device on their own (8192
// workgroup tiles and more); they run through the single-mesh pipeline, whose
// kernels are tuned for one large mesh, instead of the batched pass.
constexpr uint64_t SOLO_SLAB_VERTICES = (uint64_t)1 << 21;

// A stream extracts the timesteps it has queued in windows of up to this many
// vertices (2^27 = 256^3 x 8): the planes inside a window are classified once
// instead of once per slab, and one launch covers several slabs.
constexpr uint64_t WINDOW_VERTICES = (uint64_t)1 << 27;

struct QueuedStep {
  ftk2h_field_view views[3];
  std::vector<char> host[3];  // owned copies of host views
};

struct ServedStream {
  ftk2h_stream* s = nullptr;
  ftk2h_predicate_desc pred{};
  int ncomp = 0;
  int d = 0;
  uint64_t dims[3] = {1, 1, 1};
  uint64_t max_t = 0, appended = 0;
  uint64_t reserved = 0;
  std::deque<QueuedStep> queue;
  bool closed = false;
  bool finished = false;
  ftk2h_result* result = nullptr;
};

using BatchKey = std::tuple<int, int, int, int, double, double, double>;

BatchKey key_of(const ftk2h_predicate_desc& p, int d) {
  return BatchKey(p.kind, d, p.codimension, p.dtype, p.sos_q, p.thresholds[0], p.thresholds[1]);
}

}  // namespace

struct ftk2h_server {
  ftk2h_ctx* ctx = nullptr;
  uint64_t budget = 0, reserved = 0;
  uint64_t next_id = 1;
  std::map<uint64_t, ServedStream> streams;
  std::map<BatchKey, ftk2h_batch*> batches;
  DevBuf counts;                    // per-slab lower-plane node counts of a step
  std::vector<uint32_t> host_counts;
  hipStream_t st = nullptr;         // single-mesh slabs
  hipStream_t st_batch = nullptr;   // batched passes, concurrent with them
  // nodes per window vertex, the largest seen in a served window so far (sizes
  // the reservations of later admissions)
  double density = 0.25;
  std::mutex lock;
};

namespace {

int n_types_of(int d, int m) { return n_types_rt(d, m); }

// Whether the batched pass is built for a predicate template: star-fan cells
// (codimension 2 on 3-D + time meshes) need a global sort and are not batched, so
// their slabs run through the single-mesh pipeline.
bool batchable(const ftk2h_predicate_desc& p, int d) { return !(d == 4 && p.codimension == 2); }

// Bytes per node of a stream's payload on the device: key, lambda, union-find
// parent and root, one edge.
uint64_t bytes_per_node(const ftk2h_predicate_desc& p) {
  const uint64_t esz = p.dtype == FTK2H_F64 ? 8 : 4;
  return 8 + (uint64_t)(p.codimension + 1) * esz + 4 + 4 + 8;
}

uint64_t spatial_vertices(int sd, const uint64_t* dims) {
  uint64_t nsp = 1;
  for (int j = 0; j < sd; ++j) nsp *= dims[j];
  return nsp;
}

// Timesteps of a stream's largest window: up to WINDOW_VERTICES vertices, at least
// one slab, at most the whole stream.
int window_planes(int sd, const uint64_t* dims, uint64_t max_t) {
  const uint64_t nsp = spatial_vertices(sd, dims);
  const uint64_t fit = std::max<uint64_t>(2, WINDOW_VERTICES / nsp);
  return (int)std::max<uint64_t>(2, std::min<uint64_t>(fit, std::min<uint64_t>(max_t, 64)));
}

// Payload of one window's result at the server's measured density (nodes per
// window vertex).
uint64_t window_payload(int sd, const uint64_t* dims, const ftk2h_predicate_desc& p, int planes, double density) {
  return (uint64_t)(density * (double)planes * (double)spatial_vertices(sd, dims) * (double)bytes_per_node(p));
}

// Committed payload a stream keeps on the device before it retires it to host
// memory: one window, and at least 4 MiB, since a retirement synchronises the stream.
uint64_t retire_threshold(int sd, const uint64_t* dims, const ftk2h_predicate_desc& p, int planes, double density) {
  return std::max<uint64_t>(window_payload(sd, dims, p, planes, density), (uint64_t)4 << 20);
}

// Device bytes reserved for a stream, at the measured density:
//   * its window's timesteps and the work arrays of one window's extraction
//     (mask, pending and degenerate masks, node offset, cell count: 30 B per vertex);
//   * the open window and the window being extracted;
//   * the committed arrays: their payload is at most the retirement threshold plus
//     the window that crosses it (and at most the whole stream's payload); their
//     capacities double as they grow and start at kDevVecMinBytes each.
// A denser request than any served before can exceed its reservation (the density
// is then raised for later admissions); an allocation beyond the device's memory
// fails with E_OOM.  ftk2h_server_stats reports the bytes actually held.
uint64_t reservation(int sd, const uint64_t* dims, const ftk2h_predicate_desc& p, int ncomp, uint64_t max_t,
                     int planes, double density) {
  const uint64_t nsp = spatial_vertices(sd, dims);
  const uint64_t esz = p.dtype == FTK2H_F64 ? 8 : 4;
  const uint64_t fields = (uint64_t)planes * nsp * esz * (uint64_t)ncomp;
  if (max_t < 2) return fields;
  const uint64_t work = (uint64_t)planes * nsp * 30;
  const uint64_t win = window_payload(sd, dims, p, planes, density);
  const uint64_t whole = (uint64_t)(density * (double)max_t * (double)nsp * (double)bytes_per_node(p));
  const uint64_t committed = std::min(retire_threshold(sd, dims, p, planes, density) + win, whole);
  return fields + work + 2 * win + 2 * committed + 5 * (uint64_t)kDevVecMinBytes;
}

int finish_stream(ServedStream& S) {
  ftk2h_result* R = nullptr;
  const int rc = ftk2h_stream_finish(S.s, &R);
  ftk2h_stream_close(S.s);
  S.s = nullptr;
  S.queue.clear();
  S.finished = true;
  if (rc) return rc;
  S.result = R;
  return FTK2H_OK;
}

}  // namespace

extern "C" {

int ftk2h_server_create(ftk2h_ctx* ctx, uint64_t hbm_budget_bytes, ftk2h_server** out) {
  if (!ctx || !out) return fail(FTK2H_E_INVALID, "null argument");
  if (hbm_budget_bytes == 0) return fail(FTK2H_E_INVALID, "zero budget");
  auto s = std::make_unique<ftk2h_server>();
  s->ctx = ctx;
  s->budget = hbm_budget_bytes;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&s->st_batch, hipStreamNonBlocking));
  *out = s.release();
  return FTK2H_OK;
}

int ftk2h_server_open(ftk2h_server* s, int spatial_ndims, const uint64_t* spatial_dims,
                      const ftk2h_predicate_desc* pred, int ncomp, uint64_t max_timesteps, uint64_t* id) {
  if (!s || !spatial_dims || !pred || !id) return fail(FTK2H_E_INVALID, "null argument");
  if (spatial_ndims < 1 || spatial_ndims > 3) return fail(FTK2H_E_INVALID, "spatial dims must be 1..3");
  if (max_timesteps == 0) return fail(FTK2H_E_INVALID, "max_timesteps must be >= 1");
  std::lock_guard<std::mutex> g(s->lock);
  // the largest window whose reservation fits what is left of the budget
  int planes = window_planes(spatial_ndims, spatial_dims, max_timesteps);
  auto need_of = [&](int w) {
    return reservation(spatial_ndims, spatial_dims, *pred, ncomp, max_timesteps, w, s->density);
  };
  uint64_t need = need_of(planes);
  while (planes > 2 && s->reserved + need > s->budget) need = need_of(--planes);
  if (need_of(2) > s->budget) return fail(FTK2H_E_OVERFLOW, "request exceeds the whole device-memory budget");
  if (s->reserved + need > s->budget) return fail(FTK2H_E_AGAIN, "device-memory budget in use; retry later");
  ServedStream S;
  int rc = ftk2h_stream_open(s->ctx, spatial_ndims, spatial_dims, pred, ncomp, &S.s);
  if (rc) return rc;
  if ((rc = stream_set_window(S.s, planes))) {
    ftk2h_stream_close(S.s);
    return rc;
  }
  // finished nodes move to host memory past about one window, so that the stream's
  // device memory stays within its reservation however long it runs
  ftk2h_stream_set_retire_bytes(S.s, retire_threshold(spatial_ndims, spatial_dims, *pred, planes, s->density));
  S.pred = *pred;
  S.ncomp = ncomp;
  S.d = spatial_ndims + 1;
  for (int j = 0; j < spatial_ndims; ++j) S.dims[j] = spatial_dims[j];
  S.max_t = max_timesteps;
  S.reserved = need;
  s->reserved += need;
  *id = s->next_id++;
  s->streams.emplace(*id, std::move(S));
  return FTK2H_OK;
}

int ftk2h_server_append(ftk2h_server* s, uint64_t id, const ftk2h_field_view* comps) {
  if (!s || !comps) return fail(FTK2H_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->lock);
  auto it = s->streams.find(id);
  if (it == s->streams.end()) return fail(FTK2H_E_INVALID, "no such stream");
  ServedStream& S = it->second;
  if (S.closed || S.finished) return fail(FTK2H_E_INVALID, "stream is closed");
  if (S.appended >= S.max_t) return fail(FTK2H_E_OVERFLOW, "more timesteps than the stream was admitted for");
  QueuedStep q;
  for (int c = 0; c < S.ncomp; ++c) {
    if (!comps[c].ptr) return fail(FTK2H_E_INVALID, "null field pointer");
    q.views[c] = comps[c];
  }
  // host views are copied now (the caller may reuse its buffers); device views are
  // read by the step that consumes them
  const int sd = S.d - 1;
  uint64_t nsp = 1;
  for (int j = 0; j < sd; ++j) nsp *= S.dims[j];
  const size_t esz = S.pred.dtype == FTK2H_F64 ? 8 : 4;
  for (int c = 0; c < S.ncomp; ++c) {
    if (comps[c].on_device) continue;
    q.host[c].resize(nsp * esz);
    int64_t lc[3] = {0, 0, 0};
    const char* src = static_cast<const char*>(comps[c].ptr);
    for (uint64_t i = 0; i < nsp; ++i) {
      int64_t idx = 0;
      for (int j = 0; j < sd; ++j) idx += lc[j] * comps[c].strides[j];
      std::memcpy(&q.host[c][i * esz], src + idx * (int64_t)esz, esz);
      for (int j = 0; j < sd; ++j) {
        if (++lc[j] < (int64_t)S.dims[j]) break;
        lc[j] = 0;
      }
    }
    q.views[c].ptr = q.host[c].data();
    int64_t st = 1;
    for (int j = 0; j < 4; ++j) {
      q.views[c].strides[j] = j < sd ? st : 0;
      if (j < sd) st *= (int64_t)S.dims[j];
    }
  }
  S.queue.push_back(std::move(q));
  S.appended++;
  return FTK2H_OK;
}

int ftk2h_server_close(ftk2h_server* s, uint64_t id) {
  if (!s) return fail(FTK2H_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->lock);
  auto it = s->streams.find(id);
  if (it == s->streams.end()) return fail(FTK2H_E_INVALID, "no such stream");
  it->second.closed = true;
  return FTK2H_OK;
}

int ftk2h_server_step(ftk2h_server* s, uint64_t* done, int max_done, int* n_done) {
  if (!s || !n_done || (max_done > 0 && !done)) return fail(FTK2H_E_INVALID, "null argument");
  *n_done = 0;
  std::lock_guard<std::mutex> g(s->lock);
  HIP_TRY(hipSetDevice(s->ctx->device));
  // 1. every stream takes queued timesteps until its next slab is ready
  struct Pending {
    ServedStream* S;
    ftk2h_job job;
    ftk2h_field_view views[3];
  };
  std::map<BatchKey, std::vector<Pending>> groups;
  std::vector<Pending> solo;
  for (auto& [id, S] : s->streams) {
    if (S.finished) continue;
    // the queued timesteps fill the stream's window (a partial one when fewer are
    // queued: it is extracted now rather than waiting for more)
    while (!S.queue.empty() && !stream_slab_ready(S.s)) {
      int rc = stream_put(S.s, S.queue.front().views);
      if (rc) return rc;
      S.queue.pop_front();
    }
    if (stream_window_pending(S.s)) {
      Pending p;
      p.S = &S;
      std::memset(&p.job, 0, sizeof(p.job));
      stream_slab_desc(S.s, &p.job.mesh, p.views);
      uint64_t nv = 1;
      for (int j = 0; j < p.job.mesh.ndims; ++j) nv *= p.job.mesh.local_dims[j];
      if (nv >= SOLO_SLAB_VERTICES || !batchable(S.pred, S.d))
        solo.push_back(p);
      else
        groups[key_of(S.pred, S.d)].push_back(p);
    }
  }
  // the timestep copies were issued on each stream's own HIP stream
  for (Pending& p : solo) HIP_TRY(hipStreamSynchronize(stream_hip(p.S->s)));
  for (auto& [key, list] : groups)
    for (Pending& p : list) HIP_TRY(hipStreamSynchronize(stream_hip(p.S->s)));
  // 2b (on a worker thread and its own HIP stream): one batched pass per predicate
  // template for the small slabs, while 2a runs the slabs that fill the device
  int batch_rc = FTK2H_OK;
  std::string batch_err;
  double batch_density = 0.0;
  auto run_groups = [&]() {
    int rc = FTK2H_OK;
    if (hipSetDevice(s->ctx->device) != hipSuccess) rc = fail(FTK2H_E_HIP, "hipSetDevice");
    for (auto& [key, list] : groups) {
      if (rc) break;
      ftk2h_batch*& b = s->batches[key];
      if (!b) {
        const ServedStream& S0 = *list.front().S;
        rc = ftk2h_batch_create(s->ctx, &S0.pred, S0.d, &b);
        if (rc) break;
        ftk2h_batch_config cfg{4, 4, 0};  // persistent grid, dynamic chunks; shapes change every step
        if ((rc = ftk2h_batch_set_config(b, &cfg))) break;
      }
      std::vector<ftk2h_job> jobs(list.size());
      for (size_t i = 0; i < list.size(); ++i) {
        jobs[i] = list[i].job;
        jobs[i].comps = list[i].views;
        jobs[i].scalar = nullptr;
      }
      std::vector<ftk2h_result*> res(list.size(), nullptr);
      rc = ftk2h_batch_run(b, (int)jobs.size(), jobs.data(), res.data(), s->st_batch);
      if (rc) break;
      // lower-plane node counts of every slab: one kernel each, one copy back
      if ((rc = s->counts.ensure(list.size() * 4))) break;
      if (s->host_counts.size() < list.size()) s->host_counts.resize(list.size());
      uint32_t* dc = static_cast<uint32_t*>(s->counts.p);
      for (size_t i = 0; i < list.size(); ++i) stream_launch_lower_count(list[i].S->s, res[i], dc + i, s->st_batch);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(s->host_counts.data(), dc, list.size() * 4, hipMemcpyDeviceToHost, s->st_batch);
      if (e == hipSuccess) e = hipStreamSynchronize(s->st_batch);
      if (e != hipSuccess) {
        rc = fail(FTK2H_E_HIP, std::string("server counts: ") + hipGetErrorString(e));
        for (auto* r : res) ftk2h_result_free(r);
        break;
      }
      for (size_t i = 0; i < list.size(); ++i) {
        uint64_t sv = 1;
        for (int j = 0; j < list[i].job.mesh.ndims; ++j) sv *= list[i].job.mesh.local_dims[j];
        batch_density = std::max(batch_density, (double)res[i]->n_nodes / (double)sv);
        if (!rc) rc = stream_take_slab_n(list[i].S->s, res[i], s->host_counts[i]);
        else ftk2h_result_free(res[i]);
        res[i] = nullptr;
      }
    }
    batch_rc = rc;
    if (rc) batch_err = error_slot();
  };
  std::thread worker;
  if (!groups.empty()) worker = std::thread(run_groups);
  // 2a. slabs that fill the device on their own: the single-mesh pipeline, one by one
  int solo_rc = FTK2H_OK;
  for (Pending& p : solo) {
    ftk2h_result* R = nullptr;
    solo_rc = ftk2h_extract_regular(s->ctx, &p.job.mesh, &p.S->pred, p.views, p.S->ncomp, nullptr, nullptr, &R,
                                    s->st);
    if (solo_rc) break;
    uint64_t sv = 1;
    for (int j = 0; j < p.job.mesh.ndims; ++j) sv *= p.job.mesh.local_dims[j];
    s->density = std::max(s->density, (double)R->n_nodes / (double)sv);
    solo_rc = stream_take_slab(p.S->s, R);
    if (solo_rc) break;
  }
  const std::string solo_err = solo_rc ? error_slot() : std::string();
  if (worker.joinable()) worker.join();
  s->density = std::max(s->density, batch_density);
  if (solo_rc) return fail(solo_rc, solo_err);
  if (batch_rc) return fail(batch_rc, batch_err);
  // 3. closed and drained streams leave with their result
  for (auto& [id, S] : s->streams) {
    if (S.finished || !S.closed || !S.queue.empty() || stream_window_pending(S.s)) continue;
    int rc = finish_stream(S);
    if (rc) return rc;
    if (*n_done < max_done) done[*n_done] = id;
    ++*n_done;
  }
  return FTK2H_OK;
}

int ftk2h_server_take(ftk2h_server* s, uint64_t id, ftk2h_result** out) {
  if (!s || !out) return fail(FTK2H_E_INVALID, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> g(s->lock);
  auto it = s->streams.find(id);
  if (it == s->streams.end()) return fail(FTK2H_E_INVALID, "no such stream");
  if (!it->second.finished) return fail(FTK2H_E_AGAIN, "stream not finished");
  *out = it->second.result;
  s->reserved -= it->second.reserved;
  s->streams.erase(it);
  return FTK2H_OK;
}

int ftk2h_server_stats(const ftk2h_server* s, uint64_t* reserved_bytes, uint64_t* device_bytes, int* active,
                       int* queued_timesteps) {
  if (!s) return fail(FTK2H_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(const_cast<ftk2h_server*>(s)->lock);
  uint64_t dev = 0;
  int act = 0, q = 0;
  for (const auto& [id, S] : s->streams) {
    if (!S.finished) {
      ++act;
      dev += stream_device_bytes(S.s);
    }
    q += (int)S.queue.size();
  }
  if (reserved_bytes) *reserved_bytes = s->reserved;
  if (device_bytes) *device_bytes = dev;
  if (active) *active = act;
  if (queued_timesteps) *queued_timesteps = q;
  return FTK2H_OK;
}

void ftk2h_server_destroy(ftk2h_server* s) {
  if (!s) return;
  for (auto& [id, S] : s->streams) {
    if (S.s) ftk2h_stream_close(S.s);
    if (S.result) ftk2h_result_free(S.result);
  }
  for (auto& [k, b] : s->batches) ftk2h_batch_destroy(b);
  s->counts.release();
  if (s->st) (void)hipStreamDestroy(s->st);
  if (s->st_batch) (void)hipStreamDestroy(s->st_batch);
  delete s;
}

}  // extern "C"
