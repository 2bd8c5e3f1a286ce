// A certificate that ZeroCrossingSolver<K, double>::solve (zero_crossing.hpp:20-90)
// succeeds on given vertex values, decided without running it.
//
// CriticalPointPredicate accepts a simplex only if the solver succeeds
// (predicate.hpp:74): no pivot below 1e-12 in magnitude and every barycentric
// coordinate in [-1e-9, 1 + 1e-9].  The certificate proves both from a priori error
// bounds, so a certified simplex is accepted exactly when the solver would accept
// it; an uncertified one is handed to the solver itself.  The certificate is only
// sufficient: it never claims success for a solve that fails.
//
// Notation: v_0..v_K are the K-vectors at the simplex's vertices (doubles), A the
// K x K matrix with columns v_i - v_K, b = -v_K, so the solver computes x = A^-1 b
// and lambda = (x, 1 - sum x).  By Cramer's rule and multilinearity,
// lambda*_i = s_i / D with D = s_0 + ... + s_K, where for K = 3
//   s_0 = -det(v1,v2,v3), s_1 = det(v0,v2,v3), s_2 = -det(v0,v1,v3), s_3 = det(v0,v1,v2)
// and for K = 2
//   s_0 = det(v1,v2), s_1 = -det(v0,v2), s_2 = det(v0,v1).
// Step 1 (exact position in [0,1]): every s_i has the sign of D or is exactly 0.
//   s_i is evaluated in FP64 with error below c_K M^K (M = largest |component|;
//   c_3 = 8e-15 > 6 gamma_3 + 6 gamma_2, c_2 = 2e-15 > 2 gamma_2); s_i is exactly
//   0 when one of its vertices is the zero vector.
// Step 2 (solver error): Gaussian elimination with partial pivoting gives
//   (A + F) x^ = b with |F| <= gamma_3K |L||U| plus the rounding of A's differences
//   (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., Thm 9.4);
//   |l_ij| <= 1 and |u_ij| <= 2^(K-1) maxA (growth factor) give
//   ||F||_inf <= (K^2 2^(K-1) 1.003 gamma_3K + K u) maxA = f_K maxA with
//   f_3 = 3.7e-14, f_2 = 5.6e-15.  With ||A^-1||_inf <= r_K maxA^(K-1) / |D| (row sums of the
//   adjugate; r_3 = 6, r_2 = 2) the coordinates move by at most
//   err = r_K f_K maxA^K / |D| (times 1 + err), and lambda_K = 1 - sum by K err + 3u.
//   The test maxA^K <= kappa_K |D|_low (kappa_3 = 375, kappa_2 = 1e4) keeps that
//   below 2.5e-10 < 1e-9, so every coordinate stays in [-1e-9, 1 + 1e-9].
// Step 3 (pivots): the pivots multiply to det(A + G) ~ D and each is at most
//   2^(K-1) maxA, so each is at least |D| / (2^(K(K-1)/2) maxA^(K-1)) (1 - tiny);
//   |D|_low > 4 * 8.2e-12 maxA^2 (K = 3) or 4 * 2.1e-12 maxA (K = 2) keeps it
//   above 1e-12.
// Inputs that are not finite, or whose magnitudes leave [1e-90, 1e90], are never
// certified (the bounds assume no overflow and no loss to underflow beyond an
// absolute 1e-280).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define FTK2H_HD __host__ __device__ __forceinline__
#else
#define FTK2H_HD inline
#endif

// Error coefficients c_K of the s_i (a test builds the certificate with them set to
// zero to show that its checker notices the missing rounding bound).
#ifndef FTK2H_CERT_C3
#define FTK2H_CERT_C3 8e-15
#endif
#ifndef FTK2H_CERT_C2
#define FTK2H_CERT_C2 2e-15
#endif

namespace ftk2h {

// det(a, b, c) = a . (b x c) as the certificate evaluates it (no fused operations).
FTK2H_HD double cert_cross_dot(const double* a, const double* b, const double* c) {
  const double e0 = b[1] * c[2] - b[2] * c[1];
  const double e1 = b[2] * c[0] - b[0] * c[2];
  const double e2 = b[0] * c[1] - b[1] * c[0];
  return a[0] * e0 + a[1] * e1 + a[2] * e2;
}

template <int K>
FTK2H_HD bool solver_certified(const double (*v)[K]) {
  static_assert(K == 2 || K == 3, "certificate for K = 2, 3");
  double S = 0.0, M = 0.0, maxA = 0.0;
  bool zero[K + 1];
  for (int i = 0; i <= K; ++i) {
    bool z = true;
    for (int c = 0; c < K; ++c) {
      const double a = fabs(v[i][c]);
      S += a;
      M = fmax(M, a);
      z = z && v[i][c] == 0.0;
    }
    zero[i] = z;
  }
  if (!(S <= 1e90) || !(M >= 1e-90)) return false;  // not finite, or out of range
  for (int i = 0; i < K; ++i)
    for (int c = 0; c < K; ++c) maxA = fmax(maxA, fabs(v[i][c] - v[K][c]));
  maxA *= 1.0 + 1e-15;  // bounds the exact differences too
  double s[K + 1];
  bool z[K + 1];
  double err;
  if constexpr (K == 3) {
    const double* v0 = v[0];
    const double* v1 = v[1];
    const double* v2 = v[2];
    const double* v3 = v[3];
    s[0] = -cert_cross_dot(v1, v2, v3);
    s[1] = cert_cross_dot(v0, v2, v3);
    s[2] = -cert_cross_dot(v0, v1, v3);
    s[3] = cert_cross_dot(v2, v0, v1);  // det(v0, v1, v2)
    z[0] = zero[1] || zero[2] || zero[3];
    z[1] = zero[0] || zero[2] || zero[3];
    z[2] = zero[0] || zero[1] || zero[3];
    z[3] = zero[0] || zero[1] || zero[2];
    err = FTK2H_CERT_C3 * M * M * M + 1e-280;
  } else {
    s[0] = v[1][0] * v[2][1] - v[1][1] * v[2][0];
    s[1] = -(v[0][0] * v[2][1] - v[0][1] * v[2][0]);
    s[2] = v[0][0] * v[1][1] - v[0][1] * v[1][0];
    z[0] = zero[1] || zero[2];
    z[1] = zero[0] || zero[2];
    z[2] = zero[0] || zero[1];
    err = FTK2H_CERT_C2 * M * M + 1e-280;
  }
  bool nonneg = true, nonpos = true;
  double Dlo = 0.0;
  for (int i = 0; i <= K; ++i) {
    if (z[i]) continue;  // exactly zero
    nonneg = nonneg && s[i] > err;
    nonpos = nonpos && s[i] < -err;
    Dlo += fabs(s[i]) - err;
  }
  if (!(nonneg || nonpos) || !(Dlo > 0.0)) return false;
  Dlo *= 1.0 - 1e-15;  // rounding of the sum
  if constexpr (K == 3)
    return maxA * maxA * maxA <= 375.0 * Dlo && Dlo > 3.3e-11 * maxA * maxA;
  else
    return maxA * maxA <= 1e4 * Dlo && Dlo > 8.4e-12 * maxA;
}

}  // namespace ftk2h
