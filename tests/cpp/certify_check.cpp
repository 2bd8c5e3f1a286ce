// Checks the solver certificate (ftk2_amd/csrc/certify.hpp) against the oracle's
// restatement of ZeroCrossingSolver<K, double>::solve (zero_crossing.hpp:20-90):
// a certified simplex must be one the solver accepts.  Cases are drawn per class
// (random, origin near a face or a vertex, nearly flat, nearly parallel vectors as
// in smooth fields, zero and duplicated vertices, extreme scales, integer lattice
// values, non-finite values); prints one line per class:
//   K class cases certified solver_ok violations
// and exits 1 if any certified case is rejected by the solver.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>

#include "../../ftk2_amd/csrc/certify.hpp"
#include "../../oracle/oracle.hpp"

using Rng = std::mt19937_64;

static double unif(Rng& g, double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); }

template <int K>
static void make_case(int cls, Rng& g, double v[K + 1][K]) {
  for (int i = 0; i <= K; ++i)
    for (int c = 0; c < K; ++c) v[i][c] = unif(g, -1, 1);
  switch (cls) {
    case 0:  // random
      break;
    case 1:
    case 2: {  // origin at barycentric coordinates with some tiny (class 2: zero/negative tiny) entries
      double lam[K + 1], s = 0;
      for (int i = 0; i <= K; ++i) {
        const double e = std::pow(10.0, -unif(g, 0, 16));
        lam[i] = unif(g, 0, 1) < 0.5 ? e : unif(g, 0, 1);
        if (cls == 2 && unif(g, 0, 1) < 0.3) lam[i] = unif(g, 0, 1) < 0.5 ? 0.0 : -e;
        s += lam[i];
      }
      double p[K] = {};
      for (int i = 0; i <= K; ++i)
        for (int c = 0; c < K; ++c) p[c] += lam[i] / s * v[i][c];
      for (int i = 0; i <= K; ++i)
        for (int c = 0; c < K; ++c) v[i][c] -= p[c];
      break;
    }
    case 3: {  // nearly flat: last vertex near the affine span of the others
      const double t = std::pow(10.0, -unif(g, 0, 15));
      double w[K];
      double s = 0;
      for (int i = 0; i < K; ++i) s += (w[i] = unif(g, -1, 2));
      for (int c = 0; c < K; ++c) {
        double x = 0;
        for (int i = 0; i < K; ++i) x += w[i] / s * v[i][c];
        v[K][c] = x + t * unif(g, -1, 1);
      }
      break;
    }
    case 4: {  // nearly parallel vectors around a common base (smooth field far from a zero)
      const double d = std::pow(10.0, -unif(g, 0, 8));
      double b[K];
      for (int c = 0; c < K; ++c) b[c] = unif(g, -1, 1);
      for (int i = 0; i <= K; ++i)
        for (int c = 0; c < K; ++c) v[i][c] = b[c] + d * v[i][c];
      break;
    }
    case 5: {  // one zero vertex (sometimes two), origin inside or not
      const int i0 = (int)(g() % (K + 1));
      for (int c = 0; c < K; ++c) v[i0][c] = unif(g, 0, 1) < 0.1 ? -0.0 : 0.0;
      if (unif(g, 0, 1) < 0.2) {
        const int i1 = (int)(g() % (K + 1));
        for (int c = 0; c < K; ++c) v[i1][c] = 0.0;
      }
      break;
    }
    case 6: {  // duplicated vertex or duplicated component
      const int i0 = (int)(g() % (K + 1)), i1 = (int)(g() % (K + 1));
      if (unif(g, 0, 1) < 0.5) {
        for (int c = 0; c < K; ++c) v[i1][c] = v[i0][c];
      } else {
        const int c = (int)(g() % K);
        v[i1][c] = v[i0][c];
      }
      break;
    }
    case 7: {  // extreme scales
      const double s = std::pow(10.0, unif(g, -120, 120));
      for (int i = 0; i <= K; ++i)
        for (int c = 0; c < K; ++c) v[i][c] *= s;
      break;
    }
    case 8: {  // small integer lattice values times a power of two or ten
      const double s = unif(g, 0, 1) < 0.5 ? std::ldexp(1.0, (int)(g() % 41) - 20) : std::pow(10.0, -(double)(g() % 9));
      for (int i = 0; i <= K; ++i)
        for (int c = 0; c < K; ++c) v[i][c] = std::round(v[i][c] * 4) * s;
      break;
    }
    case 9: {  // non-finite values
      const double bad[3] = {std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::infinity(),
                             -std::numeric_limits<double>::infinity()};
      v[g() % (K + 1)][g() % K] = bad[g() % 3];
      break;
    }
    case 10: {  // origin near a vertex: one vertex tiny
      const double e = std::pow(10.0, -unif(g, 0, 18));
      const int i0 = (int)(g() % (K + 1));
      for (int c = 0; c < K; ++c) v[i0][c] *= e;
      break;
    }
  }
}

template <int K>
static int run(long n_per_class, unsigned long long seed) {
  int bad = 0;
  for (int cls = 0; cls <= 10; ++cls) {
    Rng g(seed * 1000003ULL + (unsigned long long)(cls * 17 + K));
    long cert = 0, ok = 0, viol = 0;
    for (long it = 0; it < n_per_class; ++it) {
      double v[K + 1][K];
      make_case<K>(cls, g, v);
      const bool c = ftk2h::solver_certified<K>(v);
      double lam[K + 1];
      const bool s = oracle::Solver<K, double>::solve(v, lam);
      cert += c;
      ok += s;
      if (c && !s) {
        if (viol < 5) {
          std::printf("VIOLATION K=%d class %d:", K, cls);
          for (int i = 0; i <= K; ++i)
            for (int q = 0; q < K; ++q) std::printf(" %.17g", v[i][q]);
          std::printf("\n");
        }
        ++viol;
      }
    }
    std::printf("%d %d %ld %ld %ld %ld\n", K, cls, n_per_class, cert, ok, viol);
    bad += viol != 0;
  }
  return bad;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 200000;
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
  int bad = run<2>(n, seed) + run<3>(n, seed);
  return bad ? 1 : 0;
}
