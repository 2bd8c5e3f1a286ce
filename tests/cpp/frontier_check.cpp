// Simulates a stream's retiring union-find (ftk2_amd/csrc/frontier.hpp) on the
// host: nodes arrive in layers (the lower planes of consecutive slabs), each
// commit unites random cells over the committed layer and the next one, in slot
// space exactly as the device does, and retirement runs after random commits.  The
// final track ids must equal a plain union-find over every cell (smallest global
// index of each component).  Prints counts of the paths exercised.
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "../../ftk2_amd/csrc/frontier.hpp"

static uint32_t find(std::vector<uint32_t>& p, uint32_t x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}
static void unite(std::vector<uint32_t>& p, uint32_t a, uint32_t b) {
  a = find(p, a);
  b = find(p, b);
  if (a == b) return;
  if (a > b) std::swap(a, b);
  p[b] = a;
}

int main(int argc, char** argv) {
  const int ncases = argc > 1 ? atoi(argv[1]) : 2000;
  std::mt19937_64 rng(argc > 2 ? atoll(argv[2]) : 1);
  long retirements = 0, resolved_through_moved = 0, open_max = 0, bad = 0;
  for (int c = 0; c < ncases; ++c) {
    const int nl = 1 + rng() % 12;
    const int maxw = 1 + rng() % 30;
    std::vector<uint32_t> start(nl + 1, 0);
    for (int i = 0; i < nl; ++i) start[i + 1] = start[i] + (rng() % 5 == 0 ? 0 : rng() % (maxw + 1));
    const uint32_t n = start[nl];
    const double pretire = (rng() % 4) / 3.0;
    const int cellsz = 2 + rng() % 3;
    std::vector<uint32_t> ref(n);
    std::iota(ref.begin(), ref.end(), 0u);
    ftk2h_frontier::Frontier fr;
    std::vector<uint32_t> par;  // slot space
    for (int i = 0; i < nl; ++i) {
      // initialise the slots of layer i (first commit) and layer i + 1
      const uint64_t want = (i + 1 < nl ? start[i + 2] : start[i + 1]) - fr.shift();
      while (par.size() < want) par.push_back((uint32_t)par.size());
      // cells over layer i and layer i + 1
      const uint32_t lo = start[i], hi = i + 1 < nl ? start[i + 2] : start[i + 1];
      const int ncells = hi > lo ? (int)(rng() % (2 * (hi - lo) + 1)) : 0;
      for (int k = 0; k < ncells; ++k) {
        uint32_t v[4];
        for (int q = 0; q < cellsz; ++q) v[q] = lo + rng() % (hi - lo);
        for (int q = 1; q < cellsz; ++q) {
          unite(ref, v[0], v[q]);
          unite(par, (uint32_t)(v[0] - fr.shift()), (uint32_t)(v[q] - fr.shift()));
        }
      }
      const uint64_t L = par.size() - fr.K();
      if (i + 1 < nl && std::uniform_real_distribution<double>(0, 1)(rng) < pretire) {
        std::vector<uint32_t> roots(par.size());
        for (size_t s = 0; s < par.size(); ++s) roots[s] = find(par, (uint32_t)s);
        par = fr.retire(roots, L, start[i + 1]);
        retirements++;
        open_max = std::max<long>(open_max, (long)fr.K());
        for (size_t s = 0; s < par.size(); ++s)  // every slot points at a root of smaller or equal slot
          if (par[s] > s || par[par[s]] != par[s]) bad++;
      }
    }
    std::vector<uint32_t> roots(par.size());
    for (size_t s = 0; s < par.size(); ++s) roots[s] = find(par, (uint32_t)s);
    std::vector<uint32_t> track;
    if (fr.active()) {
      track = fr.tracks(roots, par.size() - fr.K());
      for (uint32_t g = 0; g < fr.F; ++g)
        if (fr.pending[g] != track[g]) resolved_through_moved++;
    } else {
      track = roots;
    }
    if (track.size() != n) {
      printf("case %d: %zu tracks for %u nodes\n", c, track.size(), n);
      return 1;
    }
    for (uint32_t g = 0; g < n; ++g)
      if (track[g] != find(ref, g)) {
        printf("case %d: node %u track %u expected %u\n", c, g, track[g], find(ref, g));
        return 1;
      }
  }
  printf("cases %d retirements %ld resolved_through_moved %ld open_max %ld bad_parents %ld\n", ncases, retirements,
         resolved_through_moved, open_max, bad);
  return bad ? 1 : 0;
}
