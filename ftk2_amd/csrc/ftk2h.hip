// Host side of libftk2h: contexts, dispatch over (predicate, d, m, dtype), result
// objects and the C ABI declared in include/ftk2h.h.
#include "internal.hpp"

#include <cstdlib>

using namespace ftk2h;
using namespace ftk2h_internal;

namespace ftk2h_internal {
std::string& error_slot() {
  thread_local std::string s;
  return s;
}
extern template int run_pipeline<KIND_CP, 4, 3, float>(RunArgs&);
extern template int run_pipeline<KIND_CP, 4, 3, double>(RunArgs&);
extern template int run_pipeline<KIND_CP, 3, 2, float>(RunArgs&);
extern template int run_pipeline<KIND_CP, 3, 2, double>(RunArgs&);
extern template int run_pipeline<KIND_CP, 4, 2, float>(RunArgs&);
extern template int run_pipeline<KIND_CP, 4, 2, double>(RunArgs&);
extern template int run_pipeline<KIND_CP, 2, 1, float>(RunArgs&);
extern template int run_pipeline<KIND_CP, 2, 1, double>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 4, 1, float>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 4, 1, double>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 3, 1, float>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 3, 1, double>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 2, 1, float>(RunArgs&);
extern template int run_pipeline<KIND_CONTOUR, 2, 1, double>(RunArgs&);
extern template int run_pipeline<KIND_FIBER, 4, 2, float>(RunArgs&);
extern template int run_pipeline<KIND_FIBER, 4, 2, double>(RunArgs&);
extern template int run_pipeline<KIND_FIBER, 3, 2, float>(RunArgs&);
extern template int run_pipeline<KIND_FIBER, 3, 2, double>(RunArgs&);

struct Entry {
  int kind, d, m, dtype;
  RunFn fn;
};

template <int KIND, int D, int M>
constexpr Entry entry_f() {
  return {KIND, D, M, FTK2H_F32, &run_pipeline<KIND, D, M, float>};
}
template <int KIND, int D, int M>
constexpr Entry entry_d() {
  return {KIND, D, M, FTK2H_F64, &run_pipeline<KIND, D, M, double>};
}

const Entry kEntries[] = {
    // critical points
    entry_d<KIND_CP, 4, 3>(), entry_f<KIND_CP, 4, 3>(),
    entry_d<KIND_CP, 3, 2>(), entry_f<KIND_CP, 3, 2>(),
    entry_d<KIND_CP, 4, 2>(), entry_f<KIND_CP, 4, 2>(),
    entry_d<KIND_CP, 2, 1>(), entry_f<KIND_CP, 2, 1>(),
    // contour
    entry_d<KIND_CONTOUR, 4, 1>(), entry_f<KIND_CONTOUR, 4, 1>(),
    entry_d<KIND_CONTOUR, 3, 1>(), entry_f<KIND_CONTOUR, 3, 1>(),
    entry_d<KIND_CONTOUR, 2, 1>(), entry_f<KIND_CONTOUR, 2, 1>(),
    // fiber
    entry_d<KIND_FIBER, 4, 2>(), entry_f<KIND_FIBER, 4, 2>(),
    entry_d<KIND_FIBER, 3, 2>(), entry_f<KIND_FIBER, 3, 2>(),
};

RunFn find_entry(int kind, int d, int m, int dtype) {
  for (const auto& e : kEntries)
    if (e.kind == kind && e.d == d && e.m == m && e.dtype == dtype) return e.fn;
  return nullptr;
}

// Request checks shared by the extraction entry points: predicate, dtype, component
// count, a built (kind, d, m, dtype) pipeline and the mesh descriptor.
int check_request(const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred, const ftk2h_field_view* comps,
                  int ncomp, RunFn* fn) {
  const int d = mesh->ndims;
  if (d < 2 || d > 4) return fail(FTK2H_E_UNSUPPORTED, "ndims must be 2..4");
  const int m = pred->codimension;
  const int need = components_needed(pred);
  if (need < 0) return FTK2H_E_INVALID;
  if (pred->dtype != FTK2H_F32 && pred->dtype != FTK2H_F64) return fail(FTK2H_E_INVALID, "bad dtype");
  if (ncomp != need || !comps) return fail(FTK2H_E_INVALID, "wrong number of field components");
  *fn = find_entry(pred->kind, d, m, pred->dtype);
  if (!*fn)
    return fail(FTK2H_E_UNSUPPORTED, "predicate kind " + std::to_string(pred->kind) + " with d=" +
                                         std::to_string(d) + " m=" + std::to_string(m) + " not built");
  uint64_t nvert = 0;
  bool has_cells = false;
  if (int rc = check_mesh(mesh, d, &nvert, &has_cells)) return rc;
  for (int c = 0; c < ncomp; ++c)
    if (!comps[c].ptr) return fail(FTK2H_E_INVALID, "null field pointer");
  return FTK2H_OK;
}

void init_result(ftk2h_result* res, const ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh,
                 const ftk2h_predicate_desc* pred) {
  res->device = ctx->device;
  res->kind = pred->kind;
  res->d = mesh->ndims;
  res->m = pred->codimension;
  res->dtype = pred->dtype;
  for (int j = 0; j < mesh->ndims; ++j) res->gdims[j] = mesh->global_dims[j];
  res->tstride = mesh->time_id_stride;
  // ContourPredicate::extract_it stores its threshold as the scalar (predicate.hpp:150)
  if (pred->kind == FTK2H_PRED_CONTOUR) res->scalar_value = (float)pred->thresholds[0];
}

// Mesh, predicate and field fields of P.  Device views are read in place; host views
// (allow_host) are packed into a contiguous (x fastest) device copy held by the context.
int bind_request(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred,
                 const ftk2h_field_view* comps, int ncomp, const ftk2h_field_view* scalar, bool allow_host,
                 hipStream_t st, KParams& P, uint64_t* nvert_out, bool* has_cells) {
  const int d = mesh->ndims;
  uint64_t nvert = 0;
  if (int rc = check_mesh(mesh, d, &nvert, has_cells)) return rc;
  *nvert_out = nvert;
  if (!*has_cells) return FTK2H_OK;
  init_params(P, mesh, pred, nvert);
  const size_t esz = pred->dtype == FTK2H_F64 ? 8 : 4;
  auto bind = [&](const ftk2h_field_view& fv, int slot, const void** ptr, int64_t* strides) -> int {
    if (fv.on_device) {
      *ptr = fv.ptr;
      for (int j = 0; j < 4; ++j) strides[j] = j < d ? fv.strides[j] : 0;
      return FTK2H_OK;
    }
    if (!allow_host) return fail(FTK2H_E_INVALID, "this entry point reads device views only (on_device = 1)");
    std::vector<char> pack(nvert * esz);
    pack_host_view(fv, mesh, esz, nvert, pack.data());
    int rc = ctx->fields[slot].ensure(nvert * esz);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->fields[slot].p, pack.data(), nvert * esz, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    *ptr = ctx->fields[slot].p;
    packed_strides(mesh, strides);
    return FTK2H_OK;
  };
  for (int c = 0; c < ncomp; ++c) {
    int rc = bind(comps[c], c, &P.comp[c], P.cst[c]);
    if (rc) return rc;
  }
  if (scalar && scalar->ptr && pred->kind == FTK2H_PRED_CRITICAL_POINT) {
    int rc = bind(*scalar, 3, &P.scal, P.sst);
    if (rc) return rc;
  }
  finish_params(P, mesh, ncomp);
  return FTK2H_OK;
}

}  // namespace ftk2h_internal

namespace {
// Vertex ids of the node with key k: smallest vertex + chain offsets.
template <int D, int M>
void decode_keys_t(const ftk2h_result* r, const std::vector<uint64_t>& keys, uint64_t* verts) {
  constexpr auto C = build_chains<D, M>();
  uint64_t gs[4] = {0, 0, 0, 0}, s = 1;
  for (int j = 0; j < D; ++j) {
    gs[j] = s;
    s *= r->gdims[j];
  }
  if (r->tstride) gs[D - 1] = r->tstride;  // extruded layers
  for (size_t i = 0; i < keys.size(); ++i) {
    const uint64_t vmin = keys[i] / (uint64_t)C.n;
    const int t = (int)(keys[i] % (uint64_t)C.n);
    for (int k = 0; k <= M; ++k) {
      uint64_t off = 0;
      for (int j = 0; j < D; ++j)
        if ((C.s[t][k] >> j) & 1) off += gs[j];
      verts[i * (M + 1) + k] = vmin + off;
    }
  }
}
int decode_keys(const ftk2h_result* r, const std::vector<uint64_t>& keys, uint64_t* verts) {
  switch (r->d * 10 + r->m) {
    case 21: decode_keys_t<2, 1>(r, keys, verts); break;
    case 22: decode_keys_t<2, 2>(r, keys, verts); break;
    case 31: decode_keys_t<3, 1>(r, keys, verts); break;
    case 32: decode_keys_t<3, 2>(r, keys, verts); break;
    case 33: decode_keys_t<3, 3>(r, keys, verts); break;
    case 41: decode_keys_t<4, 1>(r, keys, verts); break;
    case 42: decode_keys_t<4, 2>(r, keys, verts); break;
    case 43: decode_keys_t<4, 3>(r, keys, verts); break;
    default: return fail(FTK2H_E_UNSUPPORTED, "cannot decode keys");
  }
  return FTK2H_OK;
}
}  // namespace

extern "C" {

const char* ftk2h_last_error(void) { return error_slot().c_str(); }

int ftk2h_ctx_create(int device, ftk2h_ctx** out) {
  if (!out) return fail(FTK2H_E_INVALID, "out is NULL");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(FTK2H_E_INVALID, "no such device");
  auto ctx = std::make_unique<ftk2h_ctx>();
  ctx->device = device;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  ctx->pool = std::make_shared<DevPool>();
  ctx->pool->device = device;
  if (int rc = noalloc_init(device)) return rc;  // fixed scratch array of ftk2h_extract_regular_noalloc
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  ctx->pool->cap = total_b / 8;  // idle result blocks kept for reuse
  for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
  *out = ctx.release();
  return FTK2H_OK;
}

void ftk2h_ctx_destroy(ftk2h_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->mask.release();
  ctx->pend.release();
  ctx->degm.release();
  ctx->ccount.release();
  ctx->node_off.release();
  ctx->bsum.release();
  ctx->ebsum.release();
  ctx->xedges.release();
  ctx->xcount.release();
  ctx->scan_sums.release();
  ctx->parent.release();
  if (ctx->pool) ctx->pool->trim(0);  // results still alive keep the pool
  for (auto& f : ctx->fields) f.release();
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int ftk2h_extract_regular(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred,
                          const ftk2h_field_view* comps, int ncomp, const ftk2h_field_view* scalar,
                          const ftk2h_options* opt, ftk2h_result** out, void* hip_stream) {
  if (!ctx || !mesh || !pred || !out) return fail(FTK2H_E_INVALID, "null argument");
  *out = nullptr;
  RunFn fn = nullptr;
  if (int rc = check_request(mesh, pred, comps, ncomp, &fn)) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;

  auto res = std::make_unique<ftk2h_result>();
  init_result(res.get(), ctx, mesh, pred);
  KParams P;
  uint64_t nvert = 0;
  bool has_cells = false;
  if (int rc = bind_request(ctx, mesh, pred, comps, ncomp, scalar, true, st, P, &nvert, &has_cells)) return rc;
  if (!has_cells) {  // no hypercube: the reference visits nothing
    *out = res.release();
    return FTK2H_OK;
  }
  const int64_t nb = (int64_t)((nvert + BLOCK - 1) / BLOCK);
  int rc;
  if ((rc = ctx->mask.ensure(nvert * 8))) return rc;
  if ((rc = ctx->pend.ensure(nvert * 8))) return rc;
  if ((rc = ctx->degm.ensure(nvert * 8))) return rc;
  if ((rc = ctx->ccount.ensure(nvert * 2))) return rc;
  if ((rc = ctx->node_off.ensure(nvert * 4))) return rc;
  if ((rc = ctx->bsum.ensure((nb + 1) * 4))) return rc;
  if ((rc = ctx->ebsum.ensure((nb + 1) * 4))) return rc;
  if ((rc = ctx->scan_sums.ensure((nb / SCAN_TILE + 2) * 4))) return rc;
  P.mask = static_cast<uint64_t*>(ctx->mask.p);
  P.pend = static_cast<uint64_t*>(ctx->pend.p);
  P.degm = static_cast<uint64_t*>(ctx->degm.p);
  P.ccount = static_cast<uint16_t*>(ctx->ccount.p);
  P.node_off = static_cast<uint32_t*>(ctx->node_off.p);
  P.bsum = static_cast<uint32_t*>(ctx->bsum.p);
  P.ebsum = static_cast<uint32_t*>(ctx->ebsum.p);
  if ((rc = ctx->xcount.ensure(256))) return rc;
  P.xcount = static_cast<uint32_t*>(ctx->xcount.p);

  RunArgs A{ctx, st, P, nb, opt ? opt->node_capacity : 0, opt ? opt->cell_capacity : 0, true, res.get()};
  rc = fn(A);
  if (rc) return rc;
  *out = res.release();
  return FTK2H_OK;
}

int ftk2h_result_sizes(const ftk2h_result* r, uint64_t* n_nodes, int* node_dim, int* n_conn) {
  if (!r) return fail(FTK2H_E_INVALID, "null result");
  if (n_nodes) *n_nodes = r->n_nodes;
  if (node_dim) *node_dim = r->m;
  if (n_conn) *n_conn = r->n_conn;
  return FTK2H_OK;
}

int ftk2h_result_conn_info(const ftk2h_result* r, int i, int* dim, uint64_t* n_cells) {
  if (!r || i < 0 || i >= r->n_conn) return fail(FTK2H_E_INVALID, "bad connectivity index");
  if (dim) *dim = r->conn_dim[i];
  if (n_cells) *n_cells = r->conn_cells[i];
  return FTK2H_OK;
}


int ftk2h_result_copy_keys(const ftk2h_result* r, uint64_t* keys) {
  if (!r || !keys) return fail(FTK2H_E_INVALID, "null argument");
  if (!r->n_nodes) return FTK2H_OK;
  HIP_TRY(hipSetDevice(r->device));
  HIP_TRY(hipMemcpy(keys, r->keys, r->n_nodes * 8, hipMemcpyDeviceToHost));
  return FTK2H_OK;
}

int ftk2h_result_copy_node_range(const ftk2h_result* r, uint64_t begin, uint64_t count,
                                 uint64_t* vertices, double* lambda, float* bary, float* scalar,
                                 uint64_t* track_id) {
  if (!r) return fail(FTK2H_E_INVALID, "null result");
  if (begin > r->n_nodes || count > r->n_nodes - begin)
    return fail(FTK2H_E_INVALID, "node range beyond the result");
  const uint64_t n = count;
  if (!n) return FTK2H_OK;
  HIP_TRY(hipSetDevice(r->device));
  const int nv = r->m + 1;
  if (vertices) {
    std::vector<uint64_t> keys(n);
    HIP_TRY(hipMemcpy(keys.data(), r->keys + begin, n * 8, hipMemcpyDeviceToHost));
    int rc = decode_keys(r, keys, vertices);
    if (rc) return rc;
  }
  if (lambda || bary) {
    if (r->dtype == FTK2H_F64) {
      std::vector<double> lam(n * nv);
      HIP_TRY(hipMemcpy(lam.data(), (const double*)r->lambda + begin * nv, n * nv * 8,
                        hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < n * nv; ++i) {
        if (lambda) lambda[i] = lam[i];
        if (bary) bary[i] = (float)lam[i];
      }
    } else {
      std::vector<float> lam(n * nv);
      HIP_TRY(hipMemcpy(lam.data(), (const float*)r->lambda + begin * nv, n * nv * 4,
                        hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < n * nv; ++i) {
        if (lambda) lambda[i] = (double)lam[i];
        if (bary) bary[i] = lam[i];
      }
    }
  }
  if (scalar) {
    if (r->scalar) {
      HIP_TRY(hipMemcpy(scalar, r->scalar + begin, n * 4, hipMemcpyDeviceToHost));
    } else {
      for (uint64_t i = 0; i < n; ++i) scalar[i] = r->scalar_value;  // threshold (contour) or 0
    }
  }
  if (track_id) {
    std::vector<uint32_t> tr(n);
    HIP_TRY(hipMemcpy(tr.data(), r->track + begin, n * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i) track_id[i] = tr[i];
  }
  return FTK2H_OK;
}

int ftk2h_result_copy_nodes(const ftk2h_result* r, uint64_t* vertices, double* lambda, float* bary,
                            float* scalar, uint64_t* track_id) {
  if (!r) return fail(FTK2H_E_INVALID, "null result");
  return ftk2h_result_copy_node_range(r, 0, r->n_nodes, vertices, lambda, bary, scalar, track_id);
}

int ftk2h_result_copy_conn(const ftk2h_result* r, int i, uint32_t* indices) {
  if (!r || i < 0 || i >= r->n_conn || !indices) return fail(FTK2H_E_INVALID, "bad argument");
  HIP_TRY(hipSetDevice(r->device));
  const uint64_t n = r->conn_cells[i] * (uint64_t)(r->conn_dim[i] + 1);
  if (n) HIP_TRY(hipMemcpy(indices, r->conn[i], n * 4, hipMemcpyDeviceToHost));
  return FTK2H_OK;
}

int ftk2h_result_device_view(const ftk2h_result* r, ftk2h_device_view* out) {
  if (!r || !out) return fail(FTK2H_E_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->n_nodes = r->n_nodes;
  out->node_dim = r->m;
  out->keys = r->keys;
  out->lambda = r->lambda;
  out->track_id = r->track;
  out->scalar = r->scalar;
  out->scalar_value = r->scalar_value;
  out->n_types = r->ntypes;
  for (int j = 0; j < 4; ++j) out->global_dims[j] = r->gdims[j];
  out->time_id_stride = r->tstride;
  out->n_conn = r->n_conn;
  for (int i = 0; i < 3; ++i) {
    out->conn_dim[i] = r->conn_dim[i];
    out->conn_cells[i] = r->conn_cells[i];
    out->conn[i] = r->conn[i];
  }
  return FTK2H_OK;
}

int ftk2h_result_phase_ms(const ftk2h_result* r, float* ms, int n) {
  if (!r || !ms) return fail(FTK2H_E_INVALID, "null argument");
  for (int i = 0; i < n && i < NPHASE; ++i) ms[i] = r->phase_ms[i];
  return FTK2H_OK;
}

void ftk2h_result_free(ftk2h_result* r) { delete r; }

}  // extern "C"
