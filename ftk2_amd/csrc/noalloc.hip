// ftk2h_extract_regular_noalloc (include/ftk2h.h): the extraction pipeline of
// ftk2h_extract_regular with every work array carved out of one fixed device array
// per GPU and the result written into the caller's arrays, so that no device memory
// is allocated or freed during the call.
#include "internal.hpp"

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

// The fixed work memory.  Its host-side name is only a stand-in: host code passes the
// device address obtained by hipGetSymbolAddress (per device, in noalloc_init) to the
// kernels, never the name itself.
__device__ __align__(256) unsigned char scratch[FTK2H_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
void* g_scratch[MAX_DEVICES] = {};
std::string g_scratch_error[MAX_DEVICES];

// Offsets of the per-mesh work arrays in the scratch array; the rest of it holds the
// union-find parents and cross-workgroup edges once the counts are known.
struct Layout {
  uint64_t mask, pend, degm, ccount, node_off, bsum, ebsum, sums, xcount, tail, end;
};
Layout layout(uint64_t nvert, uint64_t nb) {
  auto al = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
  Layout L;
  uint64_t o = 0;
  L.mask = o;
  o = al(o + nvert * 8);
  L.pend = o;
  o = al(o + nvert * 8);
  L.degm = o;
  o = al(o + nvert * 8);
  L.ccount = o;
  o = al(o + nvert * 2);
  L.node_off = o;
  o = al(o + nvert * 4);
  L.bsum = o;
  o = al(o + (nb + 1) * 4);
  L.ebsum = o;
  o = al(o + (nb + 1) * 4);
  L.sums = o;
  o = al(o + (nb / SCAN_TILE + 2) * 4);
  L.xcount = o;
  o = al(o + 256);
  L.tail = o;
  L.end = FTK2H_NOALLOC_SCRATCH_BYTES;
  return L;
}

}  // namespace

namespace ftk2h_internal {

int noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return FTK2H_OK;  // the entry point reports it
  std::lock_guard<std::mutex> g(g_lock);
  if (g_scratch[device]) return FTK2H_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_error[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    (void)hipGetLastError();
    return FTK2H_OK;  // contexts stay usable; ftk2h_extract_regular_noalloc fails with this message
  }
  g_scratch[device] = p;
  return FTK2H_OK;
}

}  // namespace ftk2h_internal

extern "C" int ftk2h_extract_regular_noalloc(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh,
                                             const ftk2h_predicate_desc* pred, const ftk2h_field_view* comps,
                                             int ncomp, const ftk2h_field_view* scalar,
                                             const ftk2h_output_buffers* out, uint64_t* n_nodes,
                                             uint64_t* n_cells, void* hip_stream) {
  if (!ctx || !mesh || !pred || !out || !n_nodes || !n_cells) return fail(FTK2H_E_INVALID, "null argument");
  *n_nodes = 0;
  *n_cells = 0;
  if ((out->node_capacity && (!out->keys || !out->lambda || !out->track_id)) || (out->cell_capacity && !out->cells))
    return fail(FTK2H_E_INVALID, "null output array with a nonzero capacity");
  RunFn fn = nullptr;
  if (int rc = check_request(mesh, pred, comps, ncomp, &fn)) return rc;
  if (mesh->ndims == 4 && pred->codimension == 2)
    return fail(FTK2H_E_UNSUPPORTED, "star-fan cells need a global sort and are not built without allocation");
  if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return fail(FTK2H_E_UNSUPPORTED, "device index too large");
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_lock);
    base = static_cast<char*>(g_scratch[ctx->device]);
    if (!base) return fail(FTK2H_E_HIP, "no scratch array on this device: " + g_scratch_error[ctx->device]);
  }
  std::lock_guard<std::mutex> guard(ctx->lock);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;

  ftk2h_result res;  // describes the caller's arrays; owns nothing
  init_result(&res, ctx, mesh, pred);
  KParams P;
  uint64_t nvert = 0;
  bool has_cells = false;
  if (int rc = bind_request(ctx, mesh, pred, comps, ncomp, scalar, false, st, P, &nvert, &has_cells)) return rc;
  if (!has_cells) return FTK2H_OK;
  const int64_t nb = (int64_t)((nvert + BLOCK - 1) / BLOCK);
  const Layout L = layout(nvert, (uint64_t)nb);
  if (L.tail > L.end)
    return fail(FTK2H_E_OVERFLOW, "mesh of " + std::to_string(nvert) + " vertices needs " + std::to_string(L.tail) +
                                      " bytes of work arrays; the fixed scratch array holds " +
                                      std::to_string(L.end));
  P.mask = reinterpret_cast<uint64_t*>(base + L.mask);
  P.pend = reinterpret_cast<uint64_t*>(base + L.pend);
  P.degm = reinterpret_cast<uint64_t*>(base + L.degm);
  P.ccount = reinterpret_cast<uint16_t*>(base + L.ccount);
  P.node_off = reinterpret_cast<uint32_t*>(base + L.node_off);
  P.bsum = reinterpret_cast<uint32_t*>(base + L.bsum);
  P.ebsum = reinterpret_cast<uint32_t*>(base + L.ebsum);
  P.xcount = reinterpret_cast<uint32_t*>(base + L.xcount);
  FixedOut F;
  F.keys = out->keys;
  F.lambda = out->lambda;
  F.track = out->track_id;
  F.scalar = out->scalar;
  F.cells = out->cells;
  F.node_cap = out->node_capacity;
  F.cell_cap = out->cell_capacity;
  F.tail = base + L.tail;
  F.tail_bytes = L.end - L.tail;
  F.scan_sums = reinterpret_cast<uint32_t*>(base + L.sums);
  RunArgs A{ctx, st, P, nb, 0, 0, true, &res};
  A.fixed = &F;
  const int rc = fn(A);
  *n_nodes = res.n_nodes;
  *n_cells = res.conn_cells[0];
  // the caller owns the arrays the result describes
  res.keys = nullptr;
  res.lambda = nullptr;
  res.track = nullptr;
  res.scalar = nullptr;
  for (auto& c : res.conn) c = nullptr;
  return rc;
}
