// Seeded synthetic inputs for benchmarks and tests (host only).
//
// A field is a sum of isotropic Gaussians A * exp(-|x - c(t)|^2 / (2 sigma^2)) on a
// regular grid [x, y, (z), t] (x fastest), with centres moving linearly in time:
// c(t) = c0 + u * t.  mode 0 writes the analytic gradient (one array per spatial
// axis), mode 1 the scalar.  Optional degeneracy injectors:
//   noise      i.i.d. U[-noise, noise] added to every value
//   zero_frac  fraction of vertices whose every component is set to exactly 0
//   dup_frac   fraction of values replaced by the (pre-replacement) value of the
//              same component at the -1 neighbour along a random spatial axis
// Random numbers come from a counter-based hash (splitmix64 finaliser) keyed by
// (seed, stream, counter), so results do not depend on the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
inline uint64_t h3(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return mix(mix(mix(seed) ^ (stream * 0xd1b54a32d192ed03ull)) ^ ctr);
}
inline double u01(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return (double)(h3(seed, stream, ctr) >> 11) * 0x1p-53;
}
inline double urange(uint64_t seed, uint64_t stream, uint64_t ctr, double lo, double hi) {
  return lo + (hi - lo) * u01(seed, stream, ctr);
}

struct Gauss {
  double A, sigma, c[3], u[3];
};

struct Gen {
  int sd;
  uint64_t n[3];
  int nt;
  std::vector<Gauss> g;
  // separable tables: E[t][g][axis][x], Dv[t][g][axis][x] = -(x - c)/sigma^2
  std::vector<double> E, Dv;
  uint64_t nmax = 0;
  double noise, zero_frac, dup_frac;
  uint64_t seed;
  int mode;
  size_t tix(int t, int gi, int ax, uint64_t x) const {
    return (((size_t)t * g.size() + gi) * 3 + ax) * nmax + x;
  }
  double base(uint64_t x[3], int t, int comp, uint64_t vid) const {
    // zeroed vertex
    if (zero_frac > 0 && u01(seed, 2000, vid) < zero_frac) return 0.0;
    double acc = 0.0;
    for (size_t gi = 0; gi < g.size(); ++gi) {
      double p = g[gi].A;
      for (int a = 0; a < sd; ++a) p *= E[tix(t, (int)gi, a, x[a])];
      if (mode == 0) p *= Dv[tix(t, (int)gi, comp, x[comp])];
      acc += p;
    }
    if (noise > 0) acc += urange(seed, 1000 + comp, vid, -noise, noise);
    return acc;
  }
};

}  // namespace

extern "C" {

// Fills ncomp = (mode == 0 ? sd : 1) arrays of nx*ny*(nz)*nt values (dtype 0 float,
// 1 double).  Returns 0 on success.
int ftk2synth_gaussians(int sd, const uint64_t* spatial, int nt, int ngauss, double amp_lo,
                        double amp_hi, double sigma_lo, double sigma_hi, double vel, uint64_t seed,
                        int mode, double noise, double zero_frac, double dup_frac, int subtract_mean,
                        int dtype, void* const* out) {
  if (sd < 1 || sd > 3 || nt < 1 || ngauss < 0) return 1;
  Gen G;
  G.sd = sd;
  G.nt = nt;
  G.noise = noise;
  G.zero_frac = zero_frac;
  G.dup_frac = dup_frac;
  G.seed = seed;
  G.mode = mode;
  uint64_t nsp = 1;
  for (int a = 0; a < 3; ++a) G.n[a] = a < sd ? spatial[a] : 1;
  for (int a = 0; a < sd; ++a) {
    nsp *= G.n[a];
    if (G.n[a] > G.nmax) G.nmax = G.n[a];
  }
  G.g.resize(ngauss);
  for (int i = 0; i < ngauss; ++i) {
    Gauss& q = G.g[i];
    const double mag = urange(seed, 10, 8 * i + 0, amp_lo, amp_hi);
    q.A = u01(seed, 10, 8 * i + 1) < 0.5 ? -mag : mag;
    q.sigma = urange(seed, 10, 8 * i + 2, sigma_lo, sigma_hi);
    for (int a = 0; a < 3; ++a) {
      q.c[a] = a < sd ? urange(seed, 11 + a, i, 0.0, (double)(G.n[a] - 1)) : 0.0;
      q.u[a] = a < sd ? urange(seed, 14 + a, i, -vel, vel) : 0.0;
    }
  }
  G.E.assign((size_t)nt * ngauss * 3 * G.nmax, 0.0);
  G.Dv.assign(G.E.size(), 0.0);
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < ngauss; ++i) {
      const Gauss& q = G.g[i];
      const double s2 = q.sigma * q.sigma;
      for (int a = 0; a < sd; ++a) {
        const double c = q.c[a] + q.u[a] * t;
        for (uint64_t x = 0; x < G.n[a]; ++x) {
          const double dx = (double)x - c;
          G.E[G.tix(t, i, a, x)] = std::exp(-(dx * dx) / (2.0 * s2));
          G.Dv[G.tix(t, i, a, x)] = -dx / s2;
        }
      }
    }
  const int ncomp = mode == 0 ? sd : 1;
  const uint64_t nvert = nsp * (uint64_t)nt;
  const int64_t nrows = (int64_t)(nvert / G.n[0]);
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < nrows; ++row) {
    uint64_t r = (uint64_t)row;
    uint64_t x[3] = {0, 0, 0};
    for (int a = 1; a < sd; ++a) {
      x[a] = r % G.n[a];
      r /= G.n[a];
    }
    const int t = (int)r;
    for (uint64_t xi = 0; xi < G.n[0]; ++xi) {
      x[0] = xi;
      const uint64_t vid = (uint64_t)row * G.n[0] + xi;
      for (int c = 0; c < ncomp; ++c) {
        double val;
        const bool dup = dup_frac > 0 && u01(seed, 3000 + c, vid) < dup_frac;
        int ax = 0;
        if (dup) ax = (int)(h3(seed, 4000 + c, vid) % (uint64_t)sd);
        if (dup && x[ax] > 0) {
          uint64_t y[3] = {x[0], x[1], x[2]};
          y[ax] -= 1;
          uint64_t stride = 1;
          for (int a = 0; a < ax; ++a) stride *= G.n[a];
          val = G.base(y, t, c, vid - stride);
        } else {
          val = G.base(x, t, c, vid);
        }
        if (dtype == 0)
          static_cast<float*>(out[c])[vid] = (float)val;
        else
          static_cast<double*>(out[c])[vid] = val;
      }
    }
  }
  if (subtract_mean && mode == 1) {
    // mean over all vertices, summed row by row in a fixed order
    std::vector<double> rs(nrows, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < nrows; ++row) {
      double s = 0.0;
      for (uint64_t xi = 0; xi < G.n[0]; ++xi) {
        const uint64_t vid = (uint64_t)row * G.n[0] + xi;
        s += dtype == 0 ? (double)static_cast<float*>(out[0])[vid] : static_cast<double*>(out[0])[vid];
      }
      rs[row] = s;
    }
    double tot = 0.0;
    for (int64_t row = 0; row < nrows; ++row) tot += rs[row];
    const double mean = tot / (double)nvert;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)nvert; ++v) {
      if (dtype == 0)
        static_cast<float*>(out[0])[v] = (float)((double)static_cast<float*>(out[0])[v] - mean);
      else
        static_cast<double*>(out[0])[v] -= mean;
    }
  }
  return 0;
}
}
