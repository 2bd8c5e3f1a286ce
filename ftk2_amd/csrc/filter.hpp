// FP32 sign filter for the triple product det[A; B; C] = (A x B) . C of quantized
// 3-vectors (exact integers |q| < 2^20 held in float), the level-0 SoS test of
// CriticalPointPredicate<3, T> (predicate.hpp:66-99).
//
// Error bound (u = 2^-24, M = the largest |component| of the three vectors): each
// cross component a_i b_j - a_j b_i is two rounded products and one rounded
// difference (or one product fused into the difference), so it is within
// 2.0001 u (|a_i b_j| + |a_j b_i|) <= 4.0001 u M^2 of the exact value; the dot of
// three such components with C, one rounding per product and per addition, adds at
// most gamma_3 sum |e_k c_k| <= 3.0001 u * 3 * 2.0001 M^3.  In total
// |computed - exact| <= 30.01 u M^3 < 2^-19 M^3; the bound used is 2^-18 M^3.  M^3
// <= 2^60 cannot overflow.  A value within the bound (exact zeros included) returns
// 0 and the caller decides that triple product exactly.
#pragma once

#if defined(__HIPCC__)
#define FTK2H_FHD __host__ __device__ __forceinline__
#else
#include <cmath>
#define FTK2H_FHD inline
#endif

constexpr float FTK2H_TP32_BOUND = 0x1p-18f;  // times M^3

FTK2H_FHD void cross3f(const float* A, const float* B, float* e) {
  e[0] = fmaf(A[1], B[2], -(A[2] * B[1]));
  e[1] = fmaf(A[2], B[0], -(A[0] * B[2]));
  e[2] = fmaf(A[0], B[1], -(A[1] * B[0]));
}

// e . C with the rounding the bound above assumes.
FTK2H_FHD float dot3f(const float* e, const float* C) {
  float f = e[0] * C[0];
  f = fmaf(e[1], C[1], f);
  f = fmaf(e[2], C[2], f);
  return f;
}

// +1 / -1 when (e . C) is certainly positive / negative, 0 when |e . C| <= bnd.
FTK2H_FHD int dot_signf(const float* e, const float* C, float bnd) {
  const float f = dot3f(e, C);
  return f > bnd ? 1 : (f < -bnd ? -1 : 0);
}
