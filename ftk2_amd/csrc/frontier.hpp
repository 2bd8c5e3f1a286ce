// Host bookkeeping of a stream's retiring union-find (plain C++, no HIP).
//
// Committed nodes carry global indices 0, 1, ... in key order; the track id of a
// node is the smallest global index in its component (union by smaller root).  A
// node whose slab has been committed never appears in a later cell, so once its
// slab is done only its component's root can still change.  Retirement moves such
// nodes to host memory and keeps on the device a parent array over "slots":
//   slots [0, K)      the open roots: roots (global index < F) of components that
//                     still contain a live node, ascending
//   slots [K, K + L)  the live nodes, globals [F, F + L) (slot = global - shift)
// Slot order is global order, so union by smaller slot is union by smaller global
// index and the device kernels run unchanged in slot space.  Each retired node
// keeps the root it had when it left (`pending`); an open root that stops being a
// root records where it went (`moved`).  Roots only ever move to smaller indices,
// so following `moved` from a pending root ends at the final root.
#pragma once
#include <algorithm>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace ftk2h_frontier {

struct Frontier {
  uint64_t F = 0;                  // globals below F are retired (host resident)
  std::vector<uint32_t> table;     // open roots below F, ascending (slots [0, K))
  std::vector<uint32_t> pending;   // per retired node: its root at retirement
  std::unordered_map<uint32_t, uint32_t> moved;  // former open root -> its root then

  uint64_t K() const { return table.size(); }
  // slot = global - shift for live globals (F >= K since the open roots are distinct globals below F)
  uint64_t shift() const { return F - table.size(); }
  bool active() const { return F != 0; }
  uint32_t global(uint32_t slot) const { return slot < table.size() ? table[slot] : (uint32_t)(slot + shift()); }

  // roots[s]: root slot of slot s for s in [0, K + L), live globals [F, F + L).
  // Globals [F, newF) retire.  Returns the parent array of the new slots (the new
  // open roots, then the live globals [newF, F + L)), every entry pointing at its root.
  std::vector<uint32_t> retire(const std::vector<uint32_t>& roots, uint64_t L, uint64_t newF) {
    const uint64_t k = K();
    note_moved(roots);
    for (uint64_t s = k; s < k + (newF - F); ++s) pending.push_back(global(roots[s]));
    std::vector<uint32_t> open;
    for (uint64_t s = k + (newF - F); s < k + L; ++s) {
      const uint32_t r = global(roots[s]);
      if (r < newF) open.push_back(r);
    }
    std::sort(open.begin(), open.end());
    open.erase(std::unique(open.begin(), open.end()), open.end());
    const uint64_t k2 = open.size(), shift2 = newF - k2, live2 = F + L - newF;
    std::vector<uint32_t> init(k2 + live2);
    for (uint64_t i = 0; i < k2; ++i) init[i] = (uint32_t)i;
    for (uint64_t j = 0; j < live2; ++j) {
      const uint32_t r = global(roots[k + (newF - F) + j]);
      init[k2 + j] = r < newF ? (uint32_t)(std::lower_bound(open.begin(), open.end(), r) - open.begin())
                              : (uint32_t)(r - shift2);
    }
    table.swap(open);
    F = newF;
    return init;
  }

  // Final track ids of globals [0, F + L) from the roots of the last device state.
  std::vector<uint32_t> tracks(const std::vector<uint32_t>& roots, uint64_t L) {
    note_moved(roots);
    std::vector<uint32_t> t(F + L);
    for (uint64_t g = 0; g < F; ++g) {
      uint32_t r = pending[g];
      for (auto it = moved.find(r); it != moved.end(); it = moved.find(r)) r = it->second;
      t[g] = r;
    }
    for (uint64_t j = 0; j < L; ++j) t[F + j] = global(roots[K() + j]);
    return t;
  }

 private:
  void note_moved(const std::vector<uint32_t>& roots) {
    for (uint64_t s = 0; s < K(); ++s) {
      const uint32_t r = global(roots[s]);
      if (r != table[s]) moved[table[s]] = r;
    }
  }
};

}  // namespace ftk2h_frontier
