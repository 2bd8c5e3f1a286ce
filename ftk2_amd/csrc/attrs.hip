// Attribute recording on the device (SimplicialEngine::extract_simplex,
// hguo/ftk2 include/ftk2/core/engine.hpp:1144-1201): for every node and every
// requested slot, attributes[slot] = (float) sum_i bary_i * value_i over the node's
// vertices in ascending id order, with bary_i = (float) lambda_i
// (FeatureElement::barycentric_coords[0]) and the sum in the predicate dtype T.
// value_i is one component, or the norm sqrt(sum_c v_c^2) in T of several
// (the reference's "magnitude").
#include "internal.hpp"

#include <cmath>

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

constexpr int MAX_ATTR = 16;

struct AttrDev {
  int slot, ncomp;
  const void* ptr[16];
  int64_t st[16][4];  // element strides over local coordinates
};

struct AttrParams {
  uint64_t gdims[4];
  uint64_t tstride;  // id step of the last axis when extruded, else 0
  int64_t off[4];
  uint64_t n;
  int ntypes, nattr;
  const AttrDev* attrs;
};

template <int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void k_attributes(AttrParams A, const uint64_t* keys, const T* lambda,
                                                      float* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= A.n) return;
  const uint64_t key = keys[i];
  uint64_t r = key / (uint64_t)A.ntypes;
  const int t = (int)(key - r * (uint64_t)A.ntypes);
  int64_t l[4] = {0, 0, 0, 0};
  int sd = D;
  if (A.tstride) {  // extruded: layer first, then the spatial id
    l[D - 1] = (int64_t)(r / A.tstride) - A.off[D - 1];
    r %= A.tstride;
    sd = D - 1;
  }
  for (int j = 0; j < sd; ++j) {
    const uint64_t g = r % A.gdims[j];
    r /= A.gdims[j];
    l[j] = (int64_t)g - A.off[j];
  }
  const uint32_t pc = packed_chain<D, M>(t);
  float b[M + 1];
  int corner[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) {
    b[k] = (float)lambda[i * (M + 1) + k];
    corner[k] = k == 0 ? 0 : (int)((pc >> (4 * (k - 1))) & 0xF);
  }
  for (int a = 0; a < A.nattr; ++a) {
    const AttrDev& at = A.attrs[a];
    T acc = T(0);
#pragma unroll
    for (int k = 0; k <= M; ++k) {
      T v;
      if (at.ncomp == 1) {
        int64_t e = 0;
        for (int j = 0; j < D; ++j) e += (l[j] + ((corner[k] >> j) & 1)) * at.st[0][j];
        v = static_cast<const T*>(at.ptr[0])[e];
      } else {
        T msq = T(0);
        for (int c = 0; c < at.ncomp; ++c) {
          int64_t e = 0;
          for (int j = 0; j < D; ++j) e += (l[j] + ((corner[k] >> j) & 1)) * at.st[c][j];
          const T x = static_cast<const T*>(at.ptr[c])[e];
          msq = msq + x * x;
        }
        v = sqrt(msq);
      }
      acc = acc + (T)b[k] * v;
    }
    out[i * MAX_ATTR + at.slot] = (float)acc;
  }
}

template <int D, int M, class T>
hipError_t launch(const AttrParams& A, const ftk2h_result* R, float* out, hipStream_t st) {
  const unsigned g = (unsigned)((A.n + BLOCK - 1) / BLOCK);
  k_attributes<D, M, T><<<g, BLOCK, 0, st>>>(A, R->keys, static_cast<const T*>(R->lambda), out);
  return hipGetLastError();
}

template <class T>
hipError_t dispatch(const AttrParams& A, const ftk2h_result* R, float* out, hipStream_t st) {
  switch (R->d * 10 + R->m) {
    case 21: return launch<2, 1, T>(A, R, out, st);
    case 22: return launch<2, 2, T>(A, R, out, st);
    case 31: return launch<3, 1, T>(A, R, out, st);
    case 32: return launch<3, 2, T>(A, R, out, st);
    case 33: return launch<3, 3, T>(A, R, out, st);
    case 41: return launch<4, 1, T>(A, R, out, st);
    case 42: return launch<4, 2, T>(A, R, out, st);
    case 43: return launch<4, 3, T>(A, R, out, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int ftk2h_result_add_attributes(ftk2h_ctx* ctx, ftk2h_result* R, const ftk2h_mesh_desc* mesh,
                                const ftk2h_attribute_desc* attrs, int nattr, void* hip_stream) {
  if (!ctx || !R || !mesh || (nattr > 0 && !attrs)) return fail(FTK2H_E_INVALID, "null argument");
  if (nattr < 0 || nattr > 64) return fail(FTK2H_E_INVALID, "nattr must be 0..64");
  if (mesh->ndims != R->d || mesh->time_id_stride != R->tstride)
    return fail(FTK2H_E_INVALID, "mesh differs from the result's");
  uint64_t nvert = 0;
  bool has_cells = false;
  if (int rc = check_mesh(mesh, R->d, &nvert, &has_cells)) return rc;
  for (int j = 0; j < R->d; ++j)
    if (mesh->global_dims[j] != R->gdims[j]) return fail(FTK2H_E_INVALID, "mesh differs from the result's");
  for (int a = 0; a < nattr; ++a) {
    if (attrs[a].slot < 0 || attrs[a].slot >= MAX_ATTR) return fail(FTK2H_E_INVALID, "attribute slot must be 0..15");
    if (attrs[a].ncomp < 1 || attrs[a].ncomp > 16) return fail(FTK2H_E_INVALID, "attribute ncomp must be 1..16");
    for (int c = 0; c < attrs[a].ncomp; ++c)
      if (!attrs[a].comps[c].ptr) return fail(FTK2H_E_INVALID, "null attribute view");
  }
  std::lock_guard<std::mutex> guard(ctx->lock);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
  const uint64_t n = R->n_nodes;
  if (!R->attrs && n) {
    if (R->pool)
      HIP_TRY(R->pool->get(reinterpret_cast<void**>(&R->attrs), n * MAX_ATTR * 4));
    else
      HIP_TRY(hipMalloc(&R->attrs, n * MAX_ATTR * 4));
    HIP_TRY(hipMemsetAsync(R->attrs, 0, n * MAX_ATTR * 4, st));
  }
  if (!n || nattr == 0) {
    HIP_TRY(hipStreamSynchronize(st));
    return FTK2H_OK;
  }
  // views: device ones in place, host ones packed into device copies of the local grid
  const size_t esz = R->dtype == FTK2H_F64 ? 8 : 4;
  std::vector<AttrDev> host(nattr);
  std::vector<void*> staged;
  auto cleanup = [&]() {
    for (void* p : staged) (void)hipFree(p);
  };
  for (int a = 0; a < nattr; ++a) {
    AttrDev& d = host[a];
    std::memset(&d, 0, sizeof(d));
    d.slot = attrs[a].slot;
    d.ncomp = attrs[a].ncomp;
    for (int c = 0; c < d.ncomp; ++c) {
      const ftk2h_field_view& v = attrs[a].comps[c];
      if (v.on_device) {
        d.ptr[c] = v.ptr;
        for (int j = 0; j < 4; ++j) d.st[c][j] = j < R->d ? v.strides[j] : 0;
      } else {
        std::vector<char> pack(nvert * esz);
        pack_host_view(v, mesh, esz, nvert, pack.data());
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, pack.size());
        if (e == hipSuccess) staged.push_back(p);
        if (e == hipSuccess) e = hipMemcpyAsync(p, pack.data(), pack.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
          cleanup();
          return fail(FTK2H_E_HIP, std::string("attribute copy: ") + hipGetErrorString(e));
        }
        d.ptr[c] = p;
        packed_strides(mesh, d.st[c]);
      }
    }
  }
  AttrDev* dev = nullptr;
  hipError_t e = hipMalloc(&dev, nattr * sizeof(AttrDev));
  if (e == hipSuccess) staged.push_back(dev);
  if (e == hipSuccess) e = hipMemcpyAsync(dev, host.data(), nattr * sizeof(AttrDev), hipMemcpyHostToDevice, st);
  AttrParams A{};
  for (int j = 0; j < 4; ++j) {
    A.gdims[j] = j < R->d ? R->gdims[j] : 1;
    A.off[j] = j < R->d ? (int64_t)mesh->offset[j] : 0;
  }
  A.tstride = R->tstride;
  A.n = n;
  A.ntypes = R->ntypes;
  A.nattr = nattr;
  A.attrs = dev;
  if (e == hipSuccess)
    e = R->dtype == FTK2H_F64 ? dispatch<double>(A, R, R->attrs, st) : dispatch<float>(A, R, R->attrs, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return fail(FTK2H_E_HIP, std::string("attributes: ") + hipGetErrorString(e));
  return FTK2H_OK;
}

int ftk2h_result_copy_attributes(const ftk2h_result* r, float* out) {
  if (!r || !out) return fail(FTK2H_E_INVALID, "null argument");
  const uint64_t n = r->n_nodes;
  if (!n) return FTK2H_OK;
  if (!r->attrs) {
    std::memset(out, 0, n * MAX_ATTR * sizeof(float));
    return FTK2H_OK;
  }
  HIP_TRY(hipSetDevice(r->device));
  HIP_TRY(hipMemcpy(out, r->attrs, n * MAX_ATTR * 4, hipMemcpyDeviceToHost));
  return FTK2H_OK;
}

}  // extern "C"
