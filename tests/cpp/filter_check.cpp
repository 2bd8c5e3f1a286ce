// Checks the FP32 triple-product filter (ftk2_amd/csrc/filter.hpp) against the
// exact determinant in 128-bit integers: whenever the filter returns a sign, it is
// the exact sign.  Inputs are integer 3-vectors |q| < 2^20 drawn in classes that
// reach the filter's error bound: random, C = a A + b B + small perturbation
// (nearly coplanar), nearly parallel A and B, vectors that are multiples of one
// another plus +-1, and mixed magnitudes.  Argument 3 scales the bound (a scale
// below 1 must produce violations, showing that the cases reach the bound).
// Output per class: class n certain violations.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../ftk2_amd/csrc/filter.hpp"

typedef __int128 i128;

static int exact_sign(const long long* A, const long long* B, const long long* C) {
  const i128 e0 = (i128)A[1] * B[2] - (i128)A[2] * B[1];
  const i128 e1 = (i128)A[2] * B[0] - (i128)A[0] * B[2];
  const i128 e2 = (i128)A[0] * B[1] - (i128)A[1] * B[0];
  const i128 d = e0 * C[0] + e1 * C[1] + e2 * C[2];
  return d > 0 ? 1 : (d < 0 ? -1 : 0);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  std::mt19937_64 rng(argc > 2 ? atoll(argv[2]) : 1);
  const double scale = argc > 3 ? atof(argv[3]) : 1.0;
  const long long LIM = (1 << 20) - 1;
  auto clampq = [&](double v) {
    long long q = (long long)llround(v);
    return q > LIM ? LIM : (q < -LIM ? -LIM : q);
  };
  std::uniform_real_distribution<double> U(-1, 1);
  int bad = 0;
  for (int cls = 0; cls < 5; ++cls) {
    long certain = 0, viol = 0;
    for (long it = 0; it < n; ++it) {
      long long A[3], B[3], C[3];
      const double mag = std::ldexp(1.0, 1 + (int)(rng() % 20));  // 2 .. 2^20
      for (int k = 0; k < 3; ++k) {
        A[k] = clampq(U(rng) * mag);
        B[k] = clampq(U(rng) * mag);
      }
      if (cls == 0) {
        for (int k = 0; k < 3; ++k) C[k] = clampq(U(rng) * mag);
      } else if (cls == 1) {  // nearly coplanar
        const double a = U(rng), b = U(rng), eps = std::ldexp(1.0, -(int)(rng() % 30));
        for (int k = 0; k < 3; ++k) C[k] = clampq(a * A[k] + b * B[k] + eps * U(rng) * mag);
      } else if (cls == 2) {  // nearly parallel A, B
        const double eps = std::ldexp(1.0, -(int)(rng() % 24));
        for (int k = 0; k < 3; ++k) {
          B[k] = clampq(A[k] * (1 + eps * U(rng)) + (double)(rng() % 3) - 1);
          C[k] = clampq(U(rng) * mag);
        }
      } else if (cls == 3) {  // integer multiples plus +-1
        const long long m1 = 1 + rng() % 7, m2 = 1 + rng() % 7;
        for (int k = 0; k < 3; ++k) {
          A[k] = clampq((double)A[k] / 8);
          B[k] = clampq((double)(m1 * A[k]) + (double)(rng() % 3) - 1);
          C[k] = clampq((double)(m2 * A[k] + B[k]) + (double)(rng() % 3) - 1);
        }
      } else {  // one large vector, two small ones
        for (int k = 0; k < 3; ++k) {
          A[k] = clampq(U(rng) * (1 << 20));
          B[k] = clampq(U(rng) * 4);
          C[k] = clampq(U(rng) * mag);
        }
      }
      float a[3], b[3], c[3];
      double M = 0;
      for (int k = 0; k < 3; ++k) {
        a[k] = (float)A[k];
        b[k] = (float)B[k];
        c[k] = (float)C[k];
        M = std::fmax(M, std::fmax(std::fabs((double)A[k]), std::fmax(std::fabs((double)B[k]), std::fabs((double)C[k]))));
      }
      const float bnd = (float)(M * M * M) * FTK2H_TP32_BOUND * (float)scale;
      float e[3];
      cross3f(a, b, e);
      const int s = dot_signf(e, c, bnd);
      if (s != 0) {
        certain++;
        if (s != exact_sign(A, B, C)) viol++;
      }
    }
    printf("%d %ld %ld %ld\n", cls, n, certain, viol);
    if (viol) bad = 1;
  }
  return bad;
}
