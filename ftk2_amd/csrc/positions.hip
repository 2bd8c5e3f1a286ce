// Feature positions for VTK output (hguo/ftk2 include/ftk2/utils/vtk.hpp:26-59):
// for every node, p_j = sum_i (double)bary_i * coord_j(v_i) over the node's
// vertices in ascending id order (bary_i = (float)lambda_i,
// FeatureElement::barycentric_coords[0]), summed in that order from 0.0, with the
// vertex coordinates of RegularSimplicialMesh::get_vertex_coordinates (global grid
// indices, mesh.hpp:374-378; an extruded mesh appends the layer, mesh.hpp:711-717);
// the first three coordinates, stored as float as vtkPoints stores them, and the
// "Time" array (the last coordinate on an extruded mesh, else 0).
#include "internal.hpp"

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

struct PosParams {
  uint64_t gdims[4];
  uint64_t tstride;  // id step of the last axis when extruded, else 0
  uint64_t n;
  int ntypes;
};

template <int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void k_positions(PosParams A, const uint64_t* keys, const T* lambda,
                                                     float* xyz, float* time) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= A.n) return;
  const uint64_t key = keys[i];
  uint64_t r = key / (uint64_t)A.ntypes;
  const int t = (int)(key - r * (uint64_t)A.ntypes);
  double g[4] = {0, 0, 0, 0};
  int sd = D;
  if (A.tstride) {  // extruded: layer first, then the spatial id (mesh.hpp:711-717)
    g[D - 1] = (double)(r / A.tstride);
    r %= A.tstride;
    sd = D - 1;
  }
  for (int j = 0; j < sd; ++j) {
    g[j] = (double)(r % A.gdims[j]);
    r /= A.gdims[j];
  }
  const uint32_t pc = packed_chain<D, M>(t);
  double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k <= M; ++k) {
    const int corner = k == 0 ? 0 : (int)((pc >> (4 * (k - 1))) & 0xF);
    const double b = (double)(float)lambda[i * (M + 1) + k];
#pragma unroll
    for (int j = 0; j < D; ++j) p[j] = p[j] + b * (g[j] + (double)((corner >> j) & 1));
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) xyz[i * 3 + j] = j < D ? (float)p[j] : 0.f;
  // "Time" is the last coordinate when the mesh has a layer axis (utils/vtk.hpp:54-59)
  if (time) time[i] = A.tstride ? (float)p[D - 1] : 0.f;
}

template <int D, int M, class T>
hipError_t launch(const PosParams& A, const ftk2h_result* R, float* out, float* time) {
  const unsigned g = (unsigned)((A.n + BLOCK - 1) / BLOCK);
  k_positions<D, M, T><<<g, BLOCK>>>(A, R->keys, static_cast<const T*>(R->lambda), out, time);
  return hipGetLastError();
}

template <class T>
hipError_t dispatch(const PosParams& A, const ftk2h_result* R, float* out, float* time) {
  switch (R->d * 10 + R->m) {
    case 21: return launch<2, 1, T>(A, R, out, time);
    case 22: return launch<2, 2, T>(A, R, out, time);
    case 31: return launch<3, 1, T>(A, R, out, time);
    case 32: return launch<3, 2, T>(A, R, out, time);
    case 33: return launch<3, 3, T>(A, R, out, time);
    case 41: return launch<4, 1, T>(A, R, out, time);
    case 42: return launch<4, 2, T>(A, R, out, time);
    case 43: return launch<4, 3, T>(A, R, out, time);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int ftk2h_result_copy_positions(const ftk2h_result* r, float* xyz, float* time) {
  if (!r || !xyz) return fail(FTK2H_E_INVALID, "null argument");
  const uint64_t n = r->n_nodes;
  if (!n) return FTK2H_OK;
  HIP_TRY(hipSetDevice(r->device));
  PosParams A{};
  for (int j = 0; j < 4; ++j) A.gdims[j] = j < r->d ? r->gdims[j] : 1;
  A.tstride = r->tstride;
  A.n = n;
  A.ntypes = r->ntypes;
  float* dev = nullptr;
  HIP_TRY(hipMalloc(&dev, n * 4 * sizeof(float)));  // xyz, then time
  float* dtime = dev + n * 3;
  hipError_t e = r->dtype == FTK2H_F64 ? dispatch<double>(A, r, dev, dtime) : dispatch<float>(A, r, dev, dtime);
  if (e == hipSuccess) e = hipMemcpy(xyz, dev, n * 3 * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && time) e = hipMemcpy(time, dtime, n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(FTK2H_E_HIP, std::string("positions: ") + hipGetErrorString(e));
  return FTK2H_OK;
}

}  // extern "C"
