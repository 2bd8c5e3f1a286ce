// Critical-point classification (ftk2h_result_classify, include/ftk2h.h), an opt-in
// post-pass: the reference's CriticalPointPredicate leaves FeatureElement::type at 0
// (hguo/ftk2 include/ftk2/core/predicate.hpp:80) and only its tracking configuration
// describes the classification (high_level/tracking_config.hpp:84-88).
//
// One thread per node: the spatial Jacobian J_k of the M-component field at every
// vertex k of the node's simplex (central differences over the local grid, one-sided
// at its faces, at the vertex's timestep), J = sum_k lambda_k J_k in double, then
//   gradient fields: the eigenvalue signs of the symmetric part H = (J + J^T) / 2,
//     counted exactly from its characteristic polynomial (all roots real, so the sign
//     changes of its coefficients count the positive eigenvalues, Descartes' rule);
//   flow fields: M = 1, 2 from trace / determinant / discriminant; M = 3 from the
//     Routh-Hurwitz conditions of the characteristic polynomial (all real parts
//     negative or all positive) and its discriminant (a complex pair).
#include "internal.hpp"

#include <cmath>

using namespace ftk2h;
using namespace ftk2h_internal;

namespace {

struct ClassParams {
  uint64_t gdims[4];
  int64_t off[4];
  int64_t dims[4];
  uint64_t n;
  int ntypes, mode;
  const void* comp[3];
  int64_t st[3][4];
};

template <class T>
__device__ __forceinline__ double fval(const ClassParams& C, int c, const int64_t* l, int d) {
  int64_t e = 0;
  for (int j = 0; j < d; ++j) e += l[j] * C.st[c][j];
  return (double)static_cast<const T*>(C.comp[c])[e];
}

// Number of sign changes of a coefficient sequence, zeros skipped.
template <int N>
__device__ __forceinline__ int sign_changes(const double* a) {
  int n = 0, prev = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int s = a[i] > 0 ? 1 : (a[i] < 0 ? -1 : 0);
    if (s != 0) {
      if (prev != 0 && s != prev) ++n;
      prev = s;
    }
  }
  return n;
}

template <int M>
__device__ __forceinline__ uint32_t classify_gradient(const double (*J)[M]) {
  if constexpr (M == 1) {
    return J[0][0] > 0 ? FTK2H_CP_MINIMUM : (J[0][0] < 0 ? FTK2H_CP_MAXIMUM : FTK2H_CP_DEGENERATE);
  } else {
    double H[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = 0; j < M; ++j) H[i][j] = 0.5 * (J[i][j] + J[j][i]);
    if constexpr (M == 2) {
      const double det = H[0][0] * H[1][1] - H[0][1] * H[1][0];
      const double tr = H[0][0] + H[1][1];
      if (det < 0) return FTK2H_CP_SADDLE;
      if (!(det > 0)) return FTK2H_CP_DEGENERATE;
      return tr > 0 ? FTK2H_CP_MINIMUM : FTK2H_CP_MAXIMUM;
    } else {
      // q(x) = x^3 - tr x^2 + c1 x - det; its roots are H's (real) eigenvalues
      const double tr = H[0][0] + H[1][1] + H[2][2];
      const double c1 = (H[0][0] * H[1][1] - H[0][1] * H[1][0]) + (H[0][0] * H[2][2] - H[0][2] * H[2][0]) +
                        (H[1][1] * H[2][2] - H[1][2] * H[2][1]);
      const double det = H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) -
                         H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
                         H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
      if (det == 0) return FTK2H_CP_DEGENERATE;
      const double q[4] = {1.0, -tr, c1, -det};
      const int npos = sign_changes<4>(q);
      switch (3 - npos) {  // negative eigenvalues
        case 0: return FTK2H_CP_MINIMUM;
        case 1: return FTK2H_CP_SADDLE_1;
        case 2: return FTK2H_CP_SADDLE_2;
        default: return FTK2H_CP_MAXIMUM;
      }
    }
  }
}

template <int M>
__device__ __forceinline__ uint32_t classify_flow(const double (*J)[M]) {
  if constexpr (M == 1) {
    return J[0][0] > 0 ? FTK2H_CP_SOURCE : (J[0][0] < 0 ? FTK2H_CP_SINK : FTK2H_CP_DEGENERATE);
  } else if constexpr (M == 2) {
    const double tr = J[0][0] + J[1][1];
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    if (det < 0) return FTK2H_CP_FLOW_SADDLE;
    if (!(det > 0)) return FTK2H_CP_DEGENERATE;
    const double disc = tr * tr - 4.0 * det;
    if (disc < 0) return tr > 0 ? FTK2H_CP_SPIRAL_SOURCE : (tr < 0 ? FTK2H_CP_SPIRAL_SINK : FTK2H_CP_CENTER);
    return tr > 0 ? FTK2H_CP_SOURCE : FTK2H_CP_SINK;
  } else {
    // q(x) = x^3 + a x^2 + b x + c with a = -tr, b = sum of principal 2x2 minors, c = -det
    const double tr = J[0][0] + J[1][1] + J[2][2];
    const double b = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) + (J[0][0] * J[2][2] - J[0][2] * J[2][0]) +
                     (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (det == 0) return FTK2H_CP_DEGENERATE;
    const double a = -tr, c = -det;
    const double disc = 18.0 * a * b * c - 4.0 * a * a * a * c + a * a * b * b - 4.0 * b * b * b - 27.0 * c * c;
    const bool spiral = disc < 0;  // one real root and a complex pair
    if (a > 0 && c > 0 && a * b > c) return spiral ? FTK2H_CP_SPIRAL_SINK : FTK2H_CP_SINK;      // all Re < 0
    if (a < 0 && c < 0 && a * b < c) return spiral ? FTK2H_CP_SPIRAL_SOURCE : FTK2H_CP_SOURCE;  // all Re > 0
    if (spiral && a * b == c) return FTK2H_CP_CENTER;  // the pair is purely imaginary
    return FTK2H_CP_FLOW_SADDLE;
  }
}

template <int D, int M, class T>
__global__ __launch_bounds__(BLOCK) void k_classify_cp(ClassParams C, const uint64_t* keys, const T* lambda,
                                                       uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= C.n) return;
  const uint64_t key = keys[i];
  uint64_t r = key / (uint64_t)C.ntypes;
  const int t = (int)(key - r * (uint64_t)C.ntypes);
  int64_t l0[4] = {0, 0, 0, 0};
  for (int j = 0; j < D; ++j) {
    const uint64_t g = r % C.gdims[j];
    r /= C.gdims[j];
    l0[j] = (int64_t)g - C.off[j];
  }
  const uint32_t pc = packed_chain<D, M>(t);
  double J[M][M];
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b) J[a][b] = 0.0;
#pragma unroll
  for (int k = 0; k <= M; ++k) {
    const int corner = k == 0 ? 0 : (int)((pc >> (4 * (k - 1))) & 0xF);
    int64_t l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) l[j] = l0[j] + ((corner >> j) & 1);
    const double lam = (double)lambda[i * (M + 1) + k];
#pragma unroll
    for (int ax = 0; ax < M; ++ax) {  // spatial axes are the first M
      int64_t lp[4], lm[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) lp[j] = lm[j] = l[j];
      double h = 1.0;
      if (l[ax] + 1 < C.dims[ax] && l[ax] >= 1) {
        lp[ax] += 1;
        lm[ax] -= 1;
        h = 2.0;
      } else if (l[ax] + 1 < C.dims[ax]) {
        lp[ax] += 1;
      } else if (l[ax] >= 1) {
        lm[ax] -= 1;
      } else {
        continue;  // a single vertex along this axis: no derivative
      }
#pragma unroll
      for (int c = 0; c < M; ++c) {
        const double dv = (fval<T>(C, c, lp, D) - fval<T>(C, c, lm, D)) / h;
        J[c][ax] = J[c][ax] + lam * dv;
      }
    }
  }
  out[i] = C.mode == FTK2H_CLASSIFY_GRADIENT ? classify_gradient<M>(J) : classify_flow<M>(J);
}

template <int D, int M, class T>
hipError_t launch(const ClassParams& C, const ftk2h_result* R, uint32_t* out, hipStream_t st) {
  k_classify_cp<D, M, T><<<dim3((unsigned)((C.n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st>>>(
      C, R->keys, static_cast<const T*>(R->lambda), out);
  return hipGetLastError();
}

template <class T>
hipError_t dispatch(const ClassParams& C, const ftk2h_result* R, uint32_t* out, hipStream_t st) {
  switch (R->d * 10 + R->m) {
    case 21: return launch<2, 1, T>(C, R, out, st);
    case 32: return launch<3, 2, T>(C, R, out, st);
    case 43: return launch<4, 3, T>(C, R, out, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int ftk2h_result_classify(ftk2h_ctx* ctx, ftk2h_result* R, const ftk2h_mesh_desc* mesh,
                          const ftk2h_field_view* comps, int ncomp, int mode, void* hip_stream) {
  if (!ctx || !R || !mesh || !comps) return fail(FTK2H_E_INVALID, "null argument");
  if (mode != FTK2H_CLASSIFY_GRADIENT && mode != FTK2H_CLASSIFY_FLOW) return fail(FTK2H_E_INVALID, "bad mode");
  if (R->kind != FTK2H_PRED_CRITICAL_POINT || R->m != R->d - 1)
    return fail(FTK2H_E_UNSUPPORTED, "classification needs critical points of an M-component field on M-D space + time");
  if (R->tstride) return fail(FTK2H_E_UNSUPPORTED, "classification on extruded meshes is not built");
  if (ncomp != R->m) return fail(FTK2H_E_INVALID, "wrong number of field components");
  if (mesh->ndims != R->d) return fail(FTK2H_E_INVALID, "mesh differs from the result's");
  uint64_t nvert = 0;
  bool has_cells = false;
  if (int rc = check_mesh(mesh, R->d, &nvert, &has_cells)) return rc;
  for (int j = 0; j < R->d; ++j)
    if (mesh->global_dims[j] != R->gdims[j]) return fail(FTK2H_E_INVALID, "mesh differs from the result's");
  for (int c = 0; c < ncomp; ++c)
    if (!comps[c].ptr) return fail(FTK2H_E_INVALID, "null field view");
  std::lock_guard<std::mutex> guard(ctx->lock);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
  const uint64_t n = R->n_nodes;
  if (!n) return FTK2H_OK;
  if (!R->types) {
    if (R->pool)
      HIP_TRY(R->pool->get(reinterpret_cast<void**>(&R->types), n * 4));
    else
      HIP_TRY(hipMalloc(&R->types, n * 4));
  }
  ClassParams C{};
  for (int j = 0; j < 4; ++j) {
    C.gdims[j] = j < R->d ? R->gdims[j] : 1;
    C.off[j] = j < R->d ? (int64_t)mesh->offset[j] : 0;
    C.dims[j] = j < R->d ? (int64_t)mesh->local_dims[j] : 1;
  }
  C.n = n;
  C.ntypes = R->ntypes;
  C.mode = mode;
  // views: device ones in place, host ones packed into device copies of the local grid
  const size_t esz = R->dtype == FTK2H_F64 ? 8 : 4;
  std::vector<void*> staged;
  auto cleanup = [&]() {
    for (void* p : staged) (void)hipFree(p);
  };
  for (int c = 0; c < ncomp; ++c) {
    const ftk2h_field_view& v = comps[c];
    if (v.on_device) {
      C.comp[c] = v.ptr;
      for (int j = 0; j < 4; ++j) C.st[c][j] = j < R->d ? v.strides[j] : 0;
      continue;
    }
    std::vector<char> pack(nvert * esz);
    pack_host_view(v, mesh, esz, nvert, pack.data());
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, pack.size());
    if (e == hipSuccess) staged.push_back(p);
    if (e == hipSuccess) e = hipMemcpyAsync(p, pack.data(), pack.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      cleanup();
      return fail(FTK2H_E_HIP, std::string("classification copy: ") + hipGetErrorString(e));
    }
    C.comp[c] = p;
    packed_strides(mesh, C.st[c]);
  }
  hipError_t e = R->dtype == FTK2H_F64 ? dispatch<double>(C, R, R->types, st) : dispatch<float>(C, R, R->types, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return fail(FTK2H_E_HIP, std::string("classification: ") + hipGetErrorString(e));
  return FTK2H_OK;
}

int ftk2h_result_copy_types(const ftk2h_result* r, uint32_t* types) {
  if (!r || !types) return fail(FTK2H_E_INVALID, "null argument");
  const uint64_t n = r->n_nodes;
  if (!n) return FTK2H_OK;
  if (!r->types) {
    std::memset(types, 0, n * sizeof(uint32_t));
    return FTK2H_OK;
  }
  HIP_TRY(hipSetDevice(r->device));
  HIP_TRY(hipMemcpy(types, r->types, n * 4, hipMemcpyDeviceToHost));
  return FTK2H_OK;
}

}  // extern "C"
