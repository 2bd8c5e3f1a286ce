// Streaming extraction (SimplicialEngine::execute_stream, hguo/ftk2
// include/ftk2/core/engine.hpp:214-270): timesteps arrive one at a time; each new
// timestep t closes the slab [t-1, t], a RegularSimplicialMesh with local extent 2 in
// time and offset t-1 (engine.hpp:261-266).  A stream may also extract a window of
// several slabs [t0, t0 + W] at once (ftk2h_stream_set_window; the server does so
// with the timesteps it has queued): the same mesh with local extent W + 1, so the
// planes inside the window are classified once instead of twice.  Device-resident
// state:
//   * the timesteps of the current window in `wcap` consecutive slots per field
//     component, so a window is one x-fastest array (the kernels' packed layout);
//     the next window starts with the last timestep of this one, copied to slot 0;
//   * the previous window's result: its upper-plane nodes (pure spatial simplices
//     at its last time) are also nodes of the next window, so that window stays
//     open until the next one arrives and its edges can be re-indexed by key;
//   * the committed nodes/edges and a union-find over them (the open frontier of
//     the track components).  Once the committed payload on the device exceeds the
//     retirement threshold, finished nodes, lambda and cells move to host memory and
//     the union-find keeps only the frontier and its open roots (frontier.hpp).
// The result of finish() equals the one-shot extraction over all timesteps (the
// reference's contract, test_examples.cpp:259-297), which the tests check.
#include "frontier.hpp"
#include "stream.hpp"

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

// Grow-to-fit device array that keeps its contents: capacity starts at
// kDevVecMinBytes and doubles, so it stays below twice the largest size asked for.
struct DevVec {
  void* p = nullptr;
  size_t bytes = 0, cap = 0;
  int reserve(size_t n, hipStream_t st) {
    if (n <= cap) return FTK2H_OK;
    size_t nc = cap ? cap : kDevVecMinBytes;
    while (nc < n) nc *= 2;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nc);
    if (e != hipSuccess) return fail(FTK2H_E_OOM, "stream buffer: " + std::string(hipGetErrorString(e)));
    if (bytes) {
      e = hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return fail(FTK2H_E_HIP, std::string("stream copy: ") + hipGetErrorString(e));
    }
    if (p) (void)hipFree(p);
    p = q;
    cap = nc;
    return FTK2H_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = cap = 0;
  }
};

// number of keys below `bound` (keys sorted ascending)
__global__ void k_lower_bound(const uint64_t* keys, uint64_t n, uint64_t bound, uint32_t* out) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (keys[mid] < bound)
      lo = mid + 1;
    else
      hi = mid;
  }
  *out = (uint32_t)lo;
}

// Re-index the cells of a slab result to committed indices: lower-plane nodes are
// committed at base + i; upper-plane nodes are found by key among the next slab's
// lower-plane nodes (committed at next_base + position).  Each cell is then united
// in the union-find, whose slot of a live node is its committed index - shift.
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_remap_cells(const uint32_t* cells, uint64_t nc, const uint64_t* keys,
                                                       uint32_t nlow, uint32_t base, const uint64_t* next_keys,
                                                       uint32_t next_nlow, uint32_t next_base, uint32_t* out,
                                                       uint32_t* parent, uint32_t shift, uint32_t* bad) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nc) return;
  uint32_t c[NV];
  bool ok = true;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const uint32_t x = cells[i * NV + q];
    if (x < nlow) {
      c[q] = base + x;
    } else {
      const uint64_t key = keys[x];
      uint32_t lo = 0, hi = next_nlow;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (next_keys[mid] < key)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo >= next_nlow || next_keys[lo] != key) {  // cannot happen for a consistent pair of slabs
        atomicAdd(bad, 1u);
        ok = false;
        lo = 0;
      }
      c[q] = next_base + lo;
    }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) out[i * NV + q] = c[q];
  if (!ok) return;  // reported by ftk2h_stream_finish; indices may lie beyond the parents
#pragma unroll
  for (int q = 1; q < NV; ++q) uf_unite<true>(parent, c[0] - shift, c[q] - shift);
}

__global__ __launch_bounds__(BLOCK) void k_iota(uint32_t* p, uint32_t base, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = base + (uint32_t)i;
}

}  // namespace

struct ftk2h_stream {
  ftk2h_ctx* ctx = nullptr;
  int sd = 0;
  uint64_t spatial[3] = {1, 1, 1};
  uint64_t nsp = 0;
  ftk2h_predicate_desc pred{};
  int ncomp = 0;
  size_t esz = 8;
  int ntypes = 0;
  int nv = 0;  // indices per cell (0: no connectivity)
  hipStream_t st = nullptr;
  void* fields[3] = {nullptr, nullptr, nullptr};  // wcap * nsp values each
  int wcap = 2;                                   // planes of a full window (slabs + 1)
  int np = 0;                                     // planes of the window appended so far
  uint64_t t = 0;                                 // timesteps appended
  uint64_t slabs = 0;                             // windows extracted
  ftk2h_result* open = nullptr;                   // previous slab, not committed yet
  uint32_t open_nlow = 0;
  uint32_t committed = 0;                         // committed nodes
  uint64_t committed_cells = 0;
  // device: keys/lambda of committed nodes [fr.F, committed), cells from
  // retired_cells on, parents over the union-find slots (frontier.hpp)
  DevVec keys, lambda, parent, cells, roots;
  ftk2h_frontier::Frontier fr;
  uint64_t retired_cells = 0;
  uint64_t retire_bytes = (uint64_t)8 << 30;  // device payload that triggers retirement
  std::vector<uint64_t> h_keys;               // retired payload, in committed order
  std::vector<char> h_lambda;
  std::vector<uint32_t> h_cells;
  uint32_t* scratch = nullptr;  // small device words (lower bound, bad-key count)
  float ms_total = 0.f;
  // The last committed slab: the re-indexing kernels queued on st read its arrays, so
  // they go back to the context's pool only after retiring_ev has completed.
  ftk2h_result* retiring = nullptr;
  hipEvent_t retiring_ev = nullptr;
  uint64_t node_limit = 0xFFFFFFFFull;  // committed indices are uint32 (below this)
};

namespace {

int slab_lower_count(ftk2h_stream* S, const ftk2h_result* R, uint32_t* out) {
  k_lower_bound<<<1, 1, 0, S->st>>>(R->keys, R->n_nodes, stream_lower_bound_key(S), S->scratch);
  HIP_TRY(hipMemcpyAsync(out, S->scratch, 4, hipMemcpyDeviceToHost, S->st));
  HIP_TRY(hipStreamSynchronize(S->st));
  return FTK2H_OK;
}

// Frees the result that retiring_ev guards (waiting for it first) and, when R is
// given, makes R the guarded one: an event recorded on st after the kernels that read it.
int retire_result(ftk2h_stream* S, ftk2h_result* R) {
  if (S->retiring) {
    HIP_TRY(hipEventSynchronize(S->retiring_ev));
    ftk2h_result_free(S->retiring);
    S->retiring = nullptr;
  }
  if (R) {
    S->retiring = R;
    HIP_TRY(hipEventRecord(S->retiring_ev, S->st));
  }
  return FTK2H_OK;
}

// Commit the open slab: its first `nlow` nodes (or all of them when `next` is null)
// and its cells re-indexed against `next`.
int commit_open(ftk2h_stream* S, const ftk2h_result* next, uint32_t next_nlow) {
  ftk2h_result* R = S->open;
  const uint32_t nk = next ? S->open_nlow : (uint32_t)R->n_nodes;
  // committed indices, the next slab's lower plane included, are uint32 node indices
  if ((uint64_t)S->committed + nk + (next ? next_nlow : 0) >= S->node_limit)
    return fail(FTK2H_E_OVERFLOW, "stream: more than " + std::to_string(S->node_limit - 1) +
                                      " committed nodes (uint32 indices)");
  const size_t lsz = (size_t)(R->m + 1) * S->esz;
  int rc;
  const uint64_t dev_nodes = (uint64_t)S->committed - S->fr.F;  // committed nodes still on the device
  const uint32_t shift = (uint32_t)S->fr.shift();
  if ((rc = S->keys.reserve((dev_nodes + nk) * 8, S->st))) return rc;
  if ((rc = S->lambda.reserve((dev_nodes + nk) * lsz, S->st))) return rc;
  // parents of this slab's nodes and of the next slab's lower plane (cells point there)
  const uint64_t fresh = nk + (next ? next_nlow : 0);
  const uint64_t slots = (uint64_t)S->committed - shift + fresh;
  if ((rc = S->parent.reserve(slots * 4, S->st))) return rc;
  char* kb = static_cast<char*>(S->keys.p);
  char* lb = static_cast<char*>(S->lambda.p);
  if (nk) {
    HIP_TRY(hipMemcpyAsync(kb + dev_nodes * 8, R->keys, (size_t)nk * 8, hipMemcpyDeviceToDevice, S->st));
    HIP_TRY(hipMemcpyAsync(lb + dev_nodes * lsz, R->lambda, (size_t)nk * lsz, hipMemcpyDeviceToDevice, S->st));
  }
  uint32_t* par = static_cast<uint32_t*>(S->parent.p);
  const uint64_t already = S->parent.bytes / 4;  // slots initialised so far
  if (slots > already) {
    const uint64_t cnt = slots - already;
    k_iota<<<(unsigned)((cnt + BLOCK - 1) / BLOCK), BLOCK, 0, S->st>>>(par + already, (uint32_t)already, cnt);
    S->parent.bytes = slots * 4;
  }
  if (S->nv && R->n_conn) {
    const uint64_t nc = R->conn_cells[0];
    const uint64_t dev_cells = S->committed_cells - S->retired_cells;
    if ((rc = S->cells.reserve((dev_cells + nc) * S->nv * 4, S->st))) return rc;
    uint32_t* dst = static_cast<uint32_t*>(S->cells.p) + dev_cells * S->nv;
    const unsigned g = (unsigned)((nc + BLOCK - 1) / BLOCK);
    const uint64_t* nkeys = next ? next->keys : nullptr;
    const uint32_t nb = S->committed + nk;
    switch (S->nv) {
      case 2: k_remap_cells<2><<<g, BLOCK, 0, S->st>>>(R->conn[0], nc, R->keys, nk, S->committed, nkeys, next_nlow, nb, dst, par, shift, S->scratch + 1); break;
      case 3: k_remap_cells<3><<<g, BLOCK, 0, S->st>>>(R->conn[0], nc, R->keys, nk, S->committed, nkeys, next_nlow, nb, dst, par, shift, S->scratch + 1); break;
      case 4: k_remap_cells<4><<<g, BLOCK, 0, S->st>>>(R->conn[0], nc, R->keys, nk, S->committed, nkeys, next_nlow, nb, dst, par, shift, S->scratch + 1); break;
    }
    HIP_TRY(hipGetLastError());
    S->committed_cells += nc;
    S->cells.bytes = (S->committed_cells - S->retired_cells) * S->nv * 4;
  }
  S->committed += nk;
  S->keys.bytes = ((size_t)S->committed - S->fr.F) * 8;
  S->lambda.bytes = ((size_t)S->committed - S->fr.F) * lsz;
  return FTK2H_OK;
}

// Root slot of every union-find slot, copied to host.
int slot_roots(ftk2h_stream* S, std::vector<uint32_t>& out) {
  const uint64_t P = S->parent.bytes / 4;
  out.assign(P, 0);
  if (!P) return FTK2H_OK;
  int rc;
  if ((rc = S->roots.reserve(P * 4, S->st))) return rc;
  launch_compress(static_cast<uint32_t*>(S->parent.p), static_cast<uint32_t*>(S->roots.p), P, S->st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out.data(), S->roots.p, P * 4, hipMemcpyDeviceToHost, S->st));
  HIP_TRY(hipStreamSynchronize(S->st));
  return FTK2H_OK;
}

// Move every committed node, lambda and cell on the device to host memory; the
// union-find keeps the next slab's lower plane and the open roots (frontier.hpp).
// Runs right after a commit against a next slab, so no later cell references a
// retired node.
int retire(ftk2h_stream* S) {
  const size_t lsz = (size_t)(S->pred.codimension + 1) * S->esz;
  const uint64_t nret = (uint64_t)S->committed - S->fr.F;
  const uint64_t ncell = (S->committed_cells - S->retired_cells) * S->nv;
  std::vector<uint32_t> roots;
  int rc;
  if ((rc = slot_roots(S, roots))) return rc;
  const size_t k0 = S->h_keys.size(), l0 = S->h_lambda.size(), c0 = S->h_cells.size();
  S->h_keys.resize(k0 + nret);
  S->h_lambda.resize(l0 + nret * lsz);
  S->h_cells.resize(c0 + ncell);
  if (nret) {
    HIP_TRY(hipMemcpyAsync(S->h_keys.data() + k0, S->keys.p, nret * 8, hipMemcpyDeviceToHost, S->st));
    HIP_TRY(hipMemcpyAsync(S->h_lambda.data() + l0, S->lambda.p, nret * lsz, hipMemcpyDeviceToHost, S->st));
  }
  if (ncell) HIP_TRY(hipMemcpyAsync(S->h_cells.data() + c0, S->cells.p, ncell * 4, hipMemcpyDeviceToHost, S->st));
  HIP_TRY(hipStreamSynchronize(S->st));
  const std::vector<uint32_t> init = S->fr.retire(roots, roots.size() - S->fr.K(), S->committed);
  if (!init.empty()) {
    HIP_TRY(hipMemcpyAsync(S->parent.p, init.data(), init.size() * 4, hipMemcpyHostToDevice, S->st));
    HIP_TRY(hipStreamSynchronize(S->st));
  }
  S->parent.bytes = init.size() * 4;
  S->keys.bytes = S->lambda.bytes = S->cells.bytes = 0;
  S->retired_cells = S->committed_cells;
  return FTK2H_OK;
}

}  // namespace

namespace ftk2h_internal {

int stream_put(ftk2h_stream* S, const ftk2h_field_view* comps) {
  if (stream_slab_ready(S)) return fail(FTK2H_E_INVALID, "stream: the pending window must be extracted first");
  const int64_t slot = S->np;
  // copy the timestep into its slot (x fastest, contiguous)
  for (int c = 0; c < S->ncomp; ++c) {
    char* dst = static_cast<char*>(S->fields[c]) + slot * S->nsp * S->esz;
    const ftk2h_field_view& v = comps[c];
    if (!v.ptr) return fail(FTK2H_E_INVALID, "null field pointer");
    int64_t want = 1;
    bool contiguous = true;
    for (int j = 0; j < S->sd; ++j) {
      contiguous = contiguous && (S->spatial[j] == 1 || v.strides[j] == want);
      want *= (int64_t)S->spatial[j];
    }
    if (contiguous) {
      HIP_TRY(hipMemcpyAsync(dst, v.ptr, S->nsp * S->esz, v.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             S->st));
      if (!v.on_device) HIP_TRY(hipStreamSynchronize(S->st));  // the caller may free its buffer
    } else {
      if (v.on_device) return fail(FTK2H_E_UNSUPPORTED, "stream: device views must be contiguous");
      std::vector<char> pack(S->nsp * S->esz);
      int64_t lc[3] = {0, 0, 0};
      for (uint64_t i = 0; i < S->nsp; ++i) {
        int64_t idx = 0;
        for (int j = 0; j < S->sd; ++j) idx += lc[j] * v.strides[j];
        std::memcpy(&pack[i * S->esz], static_cast<const char*>(v.ptr) + idx * (int64_t)S->esz, S->esz);
        for (int j = 0; j < S->sd; ++j) {
          if (++lc[j] < (int64_t)S->spatial[j]) break;
          lc[j] = 0;
        }
      }
      HIP_TRY(hipMemcpyAsync(dst, pack.data(), pack.size(), hipMemcpyHostToDevice, S->st));
      HIP_TRY(hipStreamSynchronize(S->st));
    }
  }
  S->np++;
  S->t++;
  return FTK2H_OK;
}

bool stream_slab_ready(const ftk2h_stream* S) { return S->np >= S->wcap; }

bool stream_window_pending(const ftk2h_stream* S) { return S->np >= 2; }

int stream_set_window(ftk2h_stream* S, int planes) {
  if (planes < 2) return fail(FTK2H_E_INVALID, "stream: a window has at least 2 timesteps");
  if (S->t) return fail(FTK2H_E_INVALID, "stream: the window is set before the first timestep");
  if (planes == S->wcap) return FTK2H_OK;
  HIP_TRY(hipSetDevice(S->ctx->device));
  for (int c = 0; c < S->ncomp; ++c) {
    if (S->fields[c]) (void)hipFree(S->fields[c]);
    S->fields[c] = nullptr;
  }
  S->wcap = planes;
  for (int c = 0; c < S->ncomp; ++c) {
    hipError_t e = hipMalloc(&S->fields[c], (size_t)planes * S->nsp * S->esz);
    if (e != hipSuccess) return fail(FTK2H_E_OOM, std::string("stream window: ") + hipGetErrorString(e));
  }
  return FTK2H_OK;
}

// The window [t - np, t - 1]: views over its slots (time stride nsp), offset t - np
// in time, global time extent t.
void stream_slab_desc(const ftk2h_stream* S, ftk2h_mesh_desc* md, ftk2h_field_view* views) {
  const uint64_t slab_t = S->t - (uint64_t)S->np;
  std::memset(md, 0, sizeof(*md));
  md->ndims = S->sd + 1;
  for (int j = 0; j < 4; ++j) {
    md->local_dims[j] = j < S->sd ? S->spatial[j] : 1;
    md->offset[j] = 0;
    md->global_dims[j] = j < S->sd ? S->spatial[j] : 1;
  }
  md->local_dims[S->sd] = (uint64_t)S->np;
  md->offset[S->sd] = slab_t;
  md->global_dims[S->sd] = S->t;
  for (int c = 0; c < S->ncomp; ++c) {
    views[c] = ftk2h_field_view{};
    views[c].on_device = 1;
    views[c].ptr = S->fields[c];
    int64_t s = 1;
    for (int j = 0; j < S->sd; ++j) {
      views[c].strides[j] = s;
      s *= (int64_t)S->spatial[j];
    }
    views[c].strides[S->sd] = (int64_t)S->nsp;
  }
}

// R = extraction of the pending slab (ownership passes to the stream): commit the
// previously open slab against it and keep R open.
int stream_take_slab(ftk2h_stream* S, ftk2h_result* R) {
  uint32_t nlow = 0;
  int rc = slab_lower_count(S, R, &nlow);
  if (rc) {
    ftk2h_result_free(R);
    return rc;
  }
  return stream_take_slab_n(S, R, nlow);
}

uint64_t stream_lower_bound_key(const ftk2h_stream* S) { return (S->t - 1) * S->nsp * (uint64_t)S->ntypes; }

void stream_launch_lower_count(const ftk2h_stream* S, const ftk2h_result* R, uint32_t* out, hipStream_t st) {
  k_lower_bound<<<1, 1, 0, st>>>(R->keys, R->n_nodes, stream_lower_bound_key(S), out);
}

// nlow = nodes of R on the slab's lower time plane.  The re-indexing kernels are
// queued on the stream's HIP stream; a missing upper-plane node is reported by
// ftk2h_stream_finish.
int stream_take_slab_n(ftk2h_stream* S, ftk2h_result* R, uint32_t nlow) {
  int rc;
  if (S->open) {
    rc = commit_open(S, R, nlow);
    const int rc2 = retire_result(S, S->open);
    S->open = nullptr;
    if (!rc) rc = rc2;
    if (!rc && S->keys.bytes + S->lambda.bytes + S->cells.bytes > S->retire_bytes) rc = retire(S);
    if (rc) {
      ftk2h_result_free(R);
      return rc;
    }
  }
  S->open = R;
  S->open_nlow = nlow;
  S->slabs++;
  // The next window starts with this one's last timestep, moved to slot 0 (the
  // extraction that read the slots has completed; a copy keeps every window one
  // ascending array, which the kernels read faster than a negative time stride).
  const int64_t last = S->np - 1;
  S->np = 1;
  for (int c = 0; c < S->ncomp; ++c) {
    char* f = static_cast<char*>(S->fields[c]);
    HIP_TRY(hipMemcpyAsync(f, f + last * S->nsp * S->esz, S->nsp * S->esz, hipMemcpyDeviceToDevice, S->st));
  }
  return FTK2H_OK;
}

uint64_t stream_timesteps(const ftk2h_stream* S) { return S->t; }

hipStream_t stream_hip(const ftk2h_stream* S) { return S->st; }

uint64_t stream_device_bytes(const ftk2h_stream* S) {
  uint64_t b = (uint64_t)S->wcap * S->nsp * S->esz * (uint64_t)S->ncomp;
  b += S->keys.cap + S->lambda.cap + S->parent.cap + S->cells.cap + S->roots.cap;
  return b;
}

}  // namespace ftk2h_internal

extern "C" {

int ftk2h_stream_open(ftk2h_ctx* ctx, int spatial_ndims, const uint64_t* spatial_dims,
                      const ftk2h_predicate_desc* pred, int ncomp, ftk2h_stream** out) {
  if (!ctx || !spatial_dims || !pred || !out) return fail(FTK2H_E_INVALID, "null argument");
  if (spatial_ndims < 1 || spatial_ndims > 3) return fail(FTK2H_E_INVALID, "spatial dims must be 1..3");
  if (ncomp < 1 || ncomp > 3) return fail(FTK2H_E_INVALID, "ncomp must be 1..3");
  auto S = std::make_unique<ftk2h_stream>();
  S->ctx = ctx;
  S->sd = spatial_ndims;
  S->nsp = 1;
  for (int j = 0; j < spatial_ndims; ++j) {
    if (spatial_dims[j] == 0) return fail(FTK2H_E_INVALID, "zero extent");
    S->spatial[j] = spatial_dims[j];
    S->nsp *= spatial_dims[j];
  }
  S->pred = *pred;
  S->ncomp = ncomp;
  S->esz = pred->dtype == FTK2H_F64 ? 8 : 4;
  const int d = spatial_ndims + 1, m = pred->codimension;
  S->ntypes = n_types_rt(d, m);
  if (!S->ntypes) return fail(FTK2H_E_UNSUPPORTED, "no simplices of this dimension");
  // cells: star edges (k=1), star-fan (k=2), marching tets/triangles (contour)
  const int k = d - m;
  if (pred->kind == FTK2H_PRED_CONTOUR && d == 4) S->nv = 4;
  else if (pred->kind == FTK2H_PRED_CONTOUR && d == 3) S->nv = 3;
  else if (m == 1 && d >= 3) S->nv = 0;
  else S->nv = k == 1 ? 2 : (k == 2 ? 3 : 0);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
  for (int c = 0; c < ncomp; ++c) HIP_TRY(hipMalloc(&S->fields[c], (size_t)S->wcap * S->nsp * S->esz));
  HIP_TRY(hipEventCreateWithFlags(&S->retiring_ev, hipEventDisableTiming));
  // FTK2H_STREAM_NODE_LIMIT lowers the committed-node limit (tests drive the overflow check)
  if (const char* lim = std::getenv("FTK2H_STREAM_NODE_LIMIT")) {
    const unsigned long long v = std::strtoull(lim, nullptr, 10);
    if (v > 0 && v < S->node_limit) S->node_limit = v;
  }
  HIP_TRY(hipMalloc(&S->scratch, 64));
  HIP_TRY(hipMemset(S->scratch, 0, 64));  // [1] counts cells whose upper node was not found
  *out = S.release();
  return FTK2H_OK;
}

// Extracts the pending window on the stream's own HIP stream and takes it.
static int extract_window(ftk2h_stream* S) {
  ftk2h_mesh_desc md;
  ftk2h_field_view views[3];
  stream_slab_desc(S, &md, views);
  ftk2h_result* R = nullptr;
  int rc = ftk2h_extract_regular(S->ctx, &md, &S->pred, views, S->ncomp, nullptr, nullptr, &R, S->st);
  if (!rc) rc = stream_take_slab(S, R);
  return rc;
}

int ftk2h_stream_append(ftk2h_stream* S, const ftk2h_field_view* comps, void* hip_stream) {
  if (!S || !comps) return fail(FTK2H_E_INVALID, "null argument");
  (void)hip_stream;  // the stream works on its own HIP stream
  HIP_TRY(hipSetDevice(S->ctx->device));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, S->st));
  int rc = stream_put(S, comps);
  if (!rc && stream_slab_ready(S)) rc = extract_window(S);
  if (!rc) {
    HIP_TRY(hipEventRecord(e1, S->st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    S->ms_total += ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int ftk2h_stream_finish(ftk2h_stream* S, ftk2h_result** out) {
  if (!S || !out) return fail(FTK2H_E_INVALID, "null argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(S->ctx->device));
  int rc;
  // a partial window (windows of more than one slab)
  if (stream_window_pending(S) && (rc = extract_window(S))) return rc;
  if (S->open) {
    rc = commit_open(S, nullptr, 0);
    const int rc2 = retire_result(S, S->open);
    S->open = nullptr;
    if (rc) return rc;
    if (rc2) return rc2;
  }
  uint32_t bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, S->scratch + 1, 4, hipMemcpyDeviceToHost, S->st));
  HIP_TRY(hipStreamSynchronize(S->st));
  if (bad) return fail(FTK2H_E_HIP, "stream: upper-plane node missing from the next slab");
  auto res = std::make_unique<ftk2h_result>();
  res->device = S->ctx->device;
  res->kind = S->pred.kind;
  res->d = S->sd + 1;
  res->m = S->pred.codimension;
  res->dtype = S->pred.dtype;
  for (int j = 0; j < S->sd; ++j) res->gdims[j] = S->spatial[j];
  res->gdims[S->sd] = S->t;
  res->ntypes = S->ntypes;
  if (S->pred.kind == FTK2H_PRED_CONTOUR) res->scalar_value = (float)S->pred.thresholds[0];
  const uint64_t n = S->committed;
  res->n_nodes = n;
  // retired payload first (host), then what is still on the device
  const uint64_t F = S->fr.F;
  std::vector<uint32_t> tracks;
  if (S->fr.active()) {
    std::vector<uint32_t> roots;
    if ((rc = slot_roots(S, roots))) return rc;
    tracks = S->fr.tracks(roots, roots.size() - S->fr.K());
    if (tracks.size() != n) return fail(FTK2H_E_HIP, "stream: retired union-find does not cover the nodes");
  }
  if (n) {
    const size_t lsz = (size_t)(res->m + 1) * S->esz;
    HIP_TRY(hipMalloc(&res->keys, n * 8));
    HIP_TRY(hipMalloc(&res->lambda, n * lsz));
    HIP_TRY(hipMalloc(&res->track, n * 4));
    if (F) {
      HIP_TRY(hipMemcpyAsync(res->keys, S->h_keys.data(), F * 8, hipMemcpyHostToDevice, S->st));
      HIP_TRY(hipMemcpyAsync(res->lambda, S->h_lambda.data(), F * lsz, hipMemcpyHostToDevice, S->st));
    }
    if (n > F) {
      HIP_TRY(hipMemcpyAsync(res->keys + F, S->keys.p, (n - F) * 8, hipMemcpyDeviceToDevice, S->st));
      HIP_TRY(hipMemcpyAsync(static_cast<char*>(res->lambda) + F * lsz, S->lambda.p, (n - F) * lsz,
                             hipMemcpyDeviceToDevice, S->st));
    }
    if (S->fr.active()) {
      HIP_TRY(hipMemcpyAsync(res->track, tracks.data(), n * 4, hipMemcpyHostToDevice, S->st));
      HIP_TRY(hipStreamSynchronize(S->st));  // `tracks` is local
    } else {
      launch_compress(static_cast<uint32_t*>(S->parent.p), res->track, n, S->st);
      HIP_TRY(hipGetLastError());
    }
  }
  if (S->nv && S->committed_cells) {
    res->n_conn = 1;
    res->conn_dim[0] = S->nv - 1;
    res->conn_cells[0] = S->committed_cells;
    const uint64_t hc = S->h_cells.size(), dc = S->committed_cells * S->nv - hc;
    HIP_TRY(hipMalloc(&res->conn[0], S->committed_cells * S->nv * 4));
    if (hc) HIP_TRY(hipMemcpyAsync(res->conn[0], S->h_cells.data(), hc * 4, hipMemcpyHostToDevice, S->st));
    if (dc) HIP_TRY(hipMemcpyAsync(res->conn[0] + hc, S->cells.p, dc * 4, hipMemcpyDeviceToDevice, S->st));
    if (S->nv == 3 && S->pred.kind != FTK2H_PRED_CONTOUR) {
      // star-fan triangles: slab order is not the global order
      if ((rc = sort_unique_cells<3>(res.get(), n, S->committed_cells, S->st))) return rc;
    }
  }
  HIP_TRY(hipStreamSynchronize(S->st));
  if ((rc = retire_result(S, nullptr))) return rc;
  res->phase_ms[0] = S->ms_total;
  *out = res.release();
  return FTK2H_OK;
}

void ftk2h_stream_close(ftk2h_stream* S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  if (S->st) (void)hipStreamSynchronize(S->st);
  if (S->open) ftk2h_result_free(S->open);
  if (S->retiring) ftk2h_result_free(S->retiring);
  if (S->retiring_ev) (void)hipEventDestroy(S->retiring_ev);
  for (auto* f : S->fields)
    if (f) (void)hipFree(f);
  if (S->scratch) (void)hipFree(S->scratch);
  S->keys.release();
  S->lambda.release();
  S->parent.release();
  S->cells.release();
  S->roots.release();
  if (S->st) (void)hipStreamDestroy(S->st);
  delete S;
}

int ftk2h_stream_set_window(ftk2h_stream* S, int timesteps) {
  if (!S) return fail(FTK2H_E_INVALID, "null argument");
  return stream_set_window(S, timesteps);
}

int ftk2h_stream_set_retire_bytes(ftk2h_stream* S, uint64_t bytes) {
  if (!S) return fail(FTK2H_E_INVALID, "null argument");
  S->retire_bytes = bytes;
  return FTK2H_OK;
}

uint64_t ftk2h_stream_resident_bytes(const ftk2h_stream* S) {
  if (!S) return 0;
  return (uint64_t)S->wcap * S->nsp * S->esz * (uint64_t)S->ncomp + S->keys.bytes + S->lambda.bytes + S->parent.bytes +
         S->cells.bytes;
}

uint64_t ftk2h_stream_retired_nodes(const ftk2h_stream* S) { return S ? S->fr.F : 0; }

}  // extern "C"
