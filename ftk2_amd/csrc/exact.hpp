// Exact sign tests and the barycentric solver, device side.
//
// Semantics follow the reference CPU code bit for bit:
//   quantization      include/ftk2/core/sos.hpp:15,25,75  (round half away from zero
//                     after a double multiply and a double add, then truncate)
//   sign              sos.hpp:12-20   (zero -> parity of the vertex id)
//   det2              sos.hpp:22-35   (zero -> +1 if i0 < i1)
//   det3              sos.hpp:72-200  (15 levels; index arguments unused; level 14 = +1)
//   origin_inside<K>  sos.hpp:202-254
//   solver            zero_crossing.hpp:20-105 (first-max partial pivoting,
//                     |pivot| < 1e-12 fails, lambda in [-1e-9, 1+1e-9])
// The library is compiled with -ffp-contract=off so no product is fused into a
// following add; fused operations appear only where written as fma() and only in
// exact-by-construction integer arithmetic or in error-bounded filters.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ftk2h {

using i128 = __int128;

template <class T>
__device__ __forceinline__ double qd(T v, double Q) {
  // double(v) * Q + (v > 0 ? 0.5 : -0.5), truncated toward zero later
  const double prod = __dmul_rn((double)v, Q);
  return __dadd_rn(prod, v > T(0) ? 0.5 : -0.5);
}

// The quantized value as an exact integer-valued double: one truncation instead
// of a conversion to a 64-bit integer.
template <class T>
__device__ __forceinline__ double qt(T v, double Q) {
  return trunc(qd(v, Q));
}

template <class T>
__device__ __forceinline__ int64_t q64(T v, double Q) {
  return (int64_t)qd(v, Q);  // truncation; inputs with |q| >= 2^62 are out of domain
}

__device__ __forceinline__ int sgn128(i128 x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

// det of [[a0,a1,1],[b0,b1,1],[c0,c1,1]]
__device__ __forceinline__ i128 det3_ones(i128 a0, i128 a1, i128 b0, i128 b1, i128 c0, i128 c1) {
  return a0 * (b1 - c1) - a1 * (b0 - c0) + (b0 * c1 - b1 * c0);
}

// The reference det3 (sos.hpp:72-200) on quantized rows X0, X1, X2; X3 = origin.
// Level 0 (the 4x4 determinant with the ones column) equals det[X0; X1; X2].
__device__ __noinline__ int det3_levels(const int64_t* X0, const int64_t* X1, const int64_t* X2) {
  const i128 a0 = X0[0], a1 = X0[1], a2 = X0[2];
  const i128 b0 = X1[0], b1 = X1[1], b2 = X1[2];
  const i128 c0 = X2[0], c1 = X2[1], c2 = X2[2];
  i128 d;
  d = a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0);  // t=0
  if (d != 0) return sgn128(d);
  d = det3_ones(b0, b1, c0, c1, 0, 0);  // t=1: rows X1,X2,X3 cols x,y
  if (d != 0) return sgn128(d);
  d = -det3_ones(b0, b2, c0, c2, 0, 0);  // t=2: cols x,z, negated
  if (d != 0) return sgn128(d);
  d = det3_ones(b1, b2, c1, c2, 0, 0);  // t=3: cols y,z
  if (d != 0) return sgn128(d);
  d = -det3_ones(a0, a1, c0, c1, 0, 0);  // t=4: rows X0,X2,X3 cols x,y, negated
  if (d != 0) return sgn128(d);
  d = c0;  // t=5: [[X2x,1],[X3x,1]]
  if (d != 0) return sgn128(d);
  d = -c1;  // t=6
  if (d != 0) return sgn128(d);
  d = det3_ones(a0, a2, c0, c2, 0, 0);  // t=7: rows X0,X2,X3 cols x,z
  if (d != 0) return sgn128(d);
  d = c2;  // t=8
  if (d != 0) return sgn128(d);
  d = -det3_ones(a1, a2, c1, c2, 0, 0);  // t=9: rows X0,X2,X3 cols y,z, negated
  if (d != 0) return sgn128(d);
  d = det3_ones(a0, a1, b0, b1, 0, 0);  // t=10: rows X0,X1,X3 cols x,y
  if (d != 0) return sgn128(d);
  d = -b0;  // t=11
  if (d != 0) return sgn128(d);
  d = b1;  // t=12
  if (d != 0) return sgn128(d);
  d = a0;  // t=13
  if (d != 0) return sgn128(d);
  return 1;  // t=14
}

// det3 levels in 64-bit integers: exact when every |q| < 2^20 (level 0 terms stay
// below 3 * 2^61; minors with the ones column below 2^43).
__device__ __forceinline__ int sgn64(int64_t x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
__device__ __forceinline__ int64_t det3_ones64(int64_t a0, int64_t a1, int64_t b0, int64_t b1) {
  // [[a0,a1,1],[b0,b1,1],[0,0,1]] (third row is the origin)
  return a0 * b1 - a1 * b0;
}
__device__ __forceinline__ int det3_levels64(const int64_t* X0, const int64_t* X1, const int64_t* X2) {
  const int64_t a0 = X0[0], a1 = X0[1], a2 = X0[2];
  const int64_t b0 = X1[0], b1 = X1[1], b2 = X1[2];
  const int64_t c0 = X2[0], c1 = X2[1], c2 = X2[2];
  int64_t d = a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0);
  if (d) return sgn64(d);
  if ((d = det3_ones64(b0, b1, c0, c1))) return sgn64(d);    // t=1
  if ((d = -det3_ones64(b0, b2, c0, c2))) return sgn64(d);   // t=2
  if ((d = det3_ones64(b1, b2, c1, c2))) return sgn64(d);    // t=3
  if ((d = -det3_ones64(a0, a1, c0, c1))) return sgn64(d);   // t=4
  if ((d = c0)) return sgn64(d);                             // t=5
  if ((d = -c1)) return sgn64(d);                            // t=6
  if ((d = det3_ones64(a0, a2, c0, c2))) return sgn64(d);    // t=7
  if ((d = c2)) return sgn64(d);                             // t=8
  if ((d = -det3_ones64(a1, a2, c1, c2))) return sgn64(d);   // t=9
  if ((d = det3_ones64(a0, a1, b0, b1))) return sgn64(d);    // t=10
  if ((d = -b0)) return sgn64(d);                            // t=11
  if ((d = b1)) return sgn64(d);                             // t=12
  if ((d = a0)) return sgn64(d);                             // t=13
  return 1;                                                  // t=14
}

// origin_inside<3> in 64-bit arithmetic (all |q| < 2^20).
__device__ __forceinline__ bool inside3_exact64(const int64_t (*X)[3]) {
  const int d0 = det3_levels64(X[1], X[3], X[2]);
  const int d1 = det3_levels64(X[0], X[2], X[3]);
  if (d0 != d1) return false;
  const int d2 = det3_levels64(X[0], X[3], X[1]);
  if (d1 != d2) return false;
  return d2 == det3_levels64(X[0], X[1], X[2]);
}

// det3 levels on integer-valued doubles with every |q| < 2^20.  Level 0 (the
// triple product, up to 6 * 2^60) is decided by an error-bounded FP64 evaluation,
// falling back to exact 64-bit integers inside the bound; levels 1-13 are minors
// below 2^43 and therefore exact in FP64.
__device__ __forceinline__ int tp_sign_d(const double* A, const double* B, const double* C, double bnd) {
  // A . (B x C); B x C is exact (products < 2^40)
  const double e0 = fma(B[1], C[2], -(B[2] * C[1]));
  const double e1 = fma(B[2], C[0], -(B[0] * C[2]));
  const double e2 = fma(B[0], C[1], -(B[1] * C[0]));
  double f = e0 * A[0];
  f = fma(e1, A[1], f);
  f = fma(e2, A[2], f);
  if (f > bnd) return 1;
  if (f < -bnd) return -1;
  // a zero row or B parallel to C (the cross product is exact) gives exactly zero
  if ((e0 == 0.0 && e1 == 0.0 && e2 == 0.0) || (A[0] == 0.0 && A[1] == 0.0 && A[2] == 0.0)) return 0;
  const int64_t s = (int64_t)e0 * (int64_t)A[0] + (int64_t)e1 * (int64_t)A[1] + (int64_t)e2 * (int64_t)A[2];
  return s > 0 ? 1 : (s < 0 ? -1 : 0);
}
__device__ __forceinline__ int sgnd(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
__device__ __forceinline__ int det3_levels_d(const double* X0, const double* X1, const double* X2, double bnd) {
  const int s0 = tp_sign_d(X0, X1, X2, bnd);
  if (s0) return s0;
  const double a0 = X0[0], a1 = X0[1], a2 = X0[2];
  const double b0 = X1[0], b1 = X1[1], b2 = X1[2];
  const double c0 = X2[0], c1 = X2[1], c2 = X2[2];
  double d;
  if ((d = fma(b0, c1, -(b1 * c0))) != 0) return sgnd(d);     // t=1
  if ((d = -fma(b0, c2, -(b2 * c0))) != 0) return sgnd(d);    // t=2
  if ((d = fma(b1, c2, -(b2 * c1))) != 0) return sgnd(d);     // t=3
  if ((d = -fma(a0, c1, -(a1 * c0))) != 0) return sgnd(d);    // t=4
  if ((d = c0) != 0) return sgnd(d);                          // t=5
  if ((d = -c1) != 0) return sgnd(d);                         // t=6
  if ((d = fma(a0, c2, -(a2 * c0))) != 0) return sgnd(d);     // t=7
  if ((d = c2) != 0) return sgnd(d);                          // t=8
  if ((d = -fma(a1, c2, -(a2 * c1))) != 0) return sgnd(d);    // t=9
  if ((d = fma(a0, b1, -(a1 * b0))) != 0) return sgnd(d);     // t=10
  if ((d = -b0) != 0) return sgnd(d);                         // t=11
  if ((d = b1) != 0) return sgnd(d);                          // t=12
  if ((d = a0) != 0) return sgnd(d);                          // t=13
  return 1;                                                   // t=14
}
// origin_inside<3> for a tet with exactly one zero vertex, on integer-valued doubles
// |q| < 2^20 (mx = max |q|): 1 inside, 0 outside, -1 when this shortcut does not
// decide it (the caller then runs inside3_exact_d).  In det3_levels_d a zero row
// makes level 0 and every level that uses the row zero, so the first level left is
// a 2x2 minor of the other two rows: level 1 (b0 c1 - b1 c0) for a zero first row,
// level 4 (-(a0 c1 - a1 c0)) for a zero second row, level 10 (a0 b1 - a1 b0) for a
// zero third row; such minors (products below 2^40) are exact in FP64, and the
// shortcut applies when all three are nonzero.  The fourth det3, the one without
// the zero row, must be decided at level 0 by its error-bounded FP64 value.  With
// m_pq = Xp.x Xq.y - Xp.y Xq.x the four dets of inside3_exact_d are, by zero vertex z:
//   d0 = det3(X1,X3,X2): z=1 -m23, z=3 -m12, z=2 +m13, z=0 -det3(X1,X2,X3)
//   d1 = det3(X0,X2,X3): z=0 +m23, z=2 -m03, z=3 +m02, z=1 +det3(X0,X2,X3)
//   d2 = det3(X0,X3,X1): z=0 -m13, z=3 -m01, z=1 +m03, z=2 -det3(X0,X1,X3)
//   d3 = det3(X0,X1,X2): z=0 +m12, z=1 -m02, z=2 +m01, z=3 +det3(X0,X1,X2)
__device__ __forceinline__ int inside3_one_zero(const double (*X)[3], double mx) {
  bool z[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) z[i] = X[i][0] == 0.0 && X[i][1] == 0.0 && X[i][2] == 0.0;
  if ((int)z[0] + (int)z[1] + (int)z[2] + (int)z[3] != 1) return -1;
  const int zi = z[0] ? 0 : (z[1] ? 1 : (z[2] ? 2 : 3));
  auto m = [&](int p, int q) { return fma(X[p][0], X[q][1], -(X[p][1] * X[q][0])); };
  const double m01 = m(0, 1), m02 = m(0, 2), m03 = m(0, 3), m12 = m(1, 2), m13 = m(1, 3), m23 = m(2, 3);
  // det3 of the three nonzero rows in ascending order, at level 0
  const double* A = zi == 0 ? X[1] : X[0];
  const double* B = zi <= 1 ? X[2] : X[1];
  const double* C = zi <= 2 ? X[3] : X[2];
  const double e0 = fma(B[1], C[2], -(B[2] * C[1]));  // exact: products below 2^40
  const double e1 = fma(B[2], C[0], -(B[0] * C[2]));
  const double e2 = fma(B[0], C[1], -(B[1] * C[0]));
  double f = e0 * A[0];
  f = fma(e1, A[1], f);
  f = fma(e2, A[2], f);
  if (!(fabs(f) > mx * mx * mx * 0x1p-48)) return -1;
  const bool pos = (zi == 0 || zi == 2) ? f < 0.0 : f > 0.0;
  const double a = zi == 0 ? m23 : (zi == 1 ? -m23 : (zi == 2 ? m13 : -m12));
  const double b = zi == 0 ? -m13 : (zi == 1 ? m03 : (zi == 2 ? -m03 : m02));
  const double c = zi == 0 ? m12 : (zi == 1 ? -m02 : (zi == 2 ? m01 : -m01));
  if (a == 0.0 || b == 0.0 || c == 0.0) return -1;
  return ((a > 0.0) == pos && (b > 0.0) == pos && (c > 0.0) == pos) ? 1 : 0;
}

// origin_inside<3> on integer-valued doubles, all |q| < 2^20 (mx = max |q|).
__device__ __forceinline__ bool inside3_exact_d(const double (*X)[3], double mx) {
  const double bnd = mx * mx * mx * 0x1p-48;
  const int d0 = det3_levels_d(X[1], X[3], X[2], bnd);
  const int d1 = det3_levels_d(X[0], X[2], X[3], bnd);
  if (d0 != d1) return false;
  const int d2 = det3_levels_d(X[0], X[3], X[1], bnd);
  if (d1 != d2) return false;
  return d2 == det3_levels_d(X[0], X[1], X[2], bnd);
}

// origin_inside<3> (sos.hpp:234-254) on quantized vertices in ascending id order.
__device__ __noinline__ bool inside3_exact(const int64_t (*X)[3]) {
  const int d0 = det3_levels(X[1], X[3], X[2]);
  const int d1 = det3_levels(X[0], X[2], X[3]);
  const int d2 = det3_levels(X[0], X[3], X[1]);
  const int d3 = det3_levels(X[0], X[1], X[2]);
  return d0 == d1 && d1 == d2 && d2 == d3;
}

// det2 (sos.hpp:22-35) for ids i0 < i1 (the tie then returns +1) or i0 > i1 (-1).
__device__ __forceinline__ int det2_exact(int64_t a, int64_t b, int64_t c, int64_t d, bool i0_lt_i1) {
  const i128 det = (i128)a * d - (i128)b * c;
  if (det > 0) return 1;
  if (det < 0) return -1;
  return i0_lt_i1 ? 1 : -1;
}

// origin_inside<2> (sos.hpp:222-232), vertices in ascending id order.
__device__ __forceinline__ bool inside2_exact(const int64_t (*X)[2]) {
  const int s0 = det2_exact(X[0][0], X[0][1], X[1][0], X[1][1], true);
  const int s1 = det2_exact(X[1][0], X[1][1], X[2][0], X[2][1], true);
  const int s2 = det2_exact(X[2][0], X[2][1], X[0][0], X[0][1], false);
  return s0 == s1 && s1 == s2;
}

// sign (sos.hpp:12-20)
template <class T>
__device__ __forceinline__ int sos_sign(T v, uint64_t id, double Q) {
  // trunc(r) > 0 <=> r >= 1; trunc(r) < 0 <=> r <= -1
  const double r = qd(v, Q);
  if (r >= 1.0) return 1;
  if (r <= -1.0) return -1;
  return (id % 2 == 0) ? 1 : -1;
}

// ZeroCrossingSolver<K,T> (zero_crossing.hpp:20-90).
template <int K, class T>
__device__ __forceinline__ bool solve_k(const T (*values)[K], T* lambda) {
  T a[K][K], x[K];
  for (int j = 0; j < K; ++j) {
    for (int i = 0; i < K; ++i) a[j][i] = values[i][j] - values[K][j];
    x[j] = -values[K][j];
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    int p = i;
#pragma unroll
    for (int j = i + 1; j < K; ++j)
      if (fabs(a[j][i]) > fabs(a[p][i])) p = j;
#pragma unroll
    for (int r = i + 1; r < K; ++r) {
      if (p == r) {
#pragma unroll
        for (int l = 0; l < K; ++l) {
          const T t = a[i][l];
          a[i][l] = a[r][l];
          a[r][l] = t;
        }
        const T t = x[i];
        x[i] = x[r];
        x[r] = t;
      }
    }
    if (fabs(a[i][i]) < T(1e-12)) return false;
#pragma unroll
    for (int j = i + 1; j < K; ++j) {
      const T f = a[j][i] / a[i][i];
      x[j] = x[j] - f * x[i];
#pragma unroll
      for (int l = i + 1; l < K; ++l) a[j][l] = a[j][l] - f * a[i][l];
    }
  }
#pragma unroll
  for (int i = K - 1; i >= 0; --i) {
    T s = T(0);
#pragma unroll
    for (int j = i + 1; j < K; ++j) s = s + a[i][j] * x[j];
    x[i] = (x[i] - s) / a[i][i];
  }
  T sum = T(0);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    lambda[i] = x[i];
    sum = sum + lambda[i];
  }
  lambda[K] = T(1.0) - sum;
  const T eps = T(1e-9);
#pragma unroll
  for (int i = 0; i <= K; ++i)
    if (lambda[i] < -eps || lambda[i] > T(1.0) + eps) return false;
  return true;
}

// ZeroCrossingSolver<1,T> (zero_crossing.hpp:93-105).
template <class T>
__device__ __forceinline__ bool solve_1(T v0, T v1, T* lambda) {
  const T d = v0 - v1;
  if (fabs(d) < T(1e-12)) return false;
  lambda[0] = -v1 / d;
  lambda[1] = T(1.0) - lambda[0];
  const T eps = T(1e-9);
  return lambda[0] >= -eps && lambda[0] <= T(1.0) + eps;
}

}  // namespace ftk2h
