// Device-wide exclusive scan and the sort + unique of connectivity cells whose
// smallest index is not owned by the emitting vertex (star-fan triangles, k = 2).
// The result is the order of FeatureComplex::sort() (complex.hpp:38-57): indices
// ascending inside a cell, cells lexicographic, duplicates removed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace ftk2h {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

// Per-tile exclusive scan; tile totals to sums[tile].
static __global__ __launch_bounds__(BLOCK) void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t* sums,
                                                             uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    run += v[i];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(run, &tot);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < n) out[base + i] = pre;
    pre += v[i];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

static __global__ __launch_bounds__(BLOCK) void k_scan_add(uint32_t* out, const uint32_t* offs, uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  const uint32_t o = offs[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    if (base + i < n) out[base + i] += o;
}

// out[0..n) = exclusive scan of in[0..n); out[n] = total.  sums: >= tiles + 1 words.
inline hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* sums, hipStream_t st) {
  const uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 0) return hipMemcpyAsync(out, sums, 0, hipMemcpyDeviceToDevice, st);
  k_scan_tiles<<<dim3((unsigned)tiles), dim3(BLOCK), 0, st>>>(in, out, sums, n);
  k_scan_small<<<1, 1024, 0, st>>>(sums, (int64_t)tiles);
  k_scan_add<<<dim3((unsigned)tiles), dim3(BLOCK), 0, st>>>(out, sums, n);
  // total: sums[tiles] after the small scan
  return hipMemcpyAsync(out + n, sums + tiles, 4, hipMemcpyDeviceToDevice, st);
}

template <int NV>
__global__ __launch_bounds__(BLOCK) void k_cell_count(const uint32_t* cells, uint64_t nc, uint32_t* cnt) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < nc) atomicAdd(cnt + cells[i * NV], 1u);
}

template <int NV>
__global__ __launch_bounds__(BLOCK) void k_cell_scatter(const uint32_t* cells, uint64_t nc, uint32_t* cursor,
                                                        uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nc) return;
  const uint32_t pos = atomicAdd(cursor + cells[i * NV], 1u);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[(uint64_t)pos * NV + q] = cells[i * NV + q];
}

// One thread per bucket (smallest index a): sort the bucket's cells and flag the
// first cell of every run of equal cells.
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_cell_bucket_sort(uint32_t* cells, const uint32_t* off, uint64_t nb,
                                                            uint32_t* keep) {
  const uint64_t a = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (a >= nb) return;
  const uint32_t b = off[a], e = off[a + 1];
  if (e == b) return;
  sort_cells<NV>(cells + (uint64_t)b * NV, e - b);
  keep[b] = 1;
  for (uint32_t i = b + 1; i < e; ++i) {
    bool same = true;
#pragma unroll
    for (int q = 0; q < NV; ++q) same = same && cells[(uint64_t)i * NV + q] == cells[(uint64_t)(i - 1) * NV + q];
    keep[i] = same ? 0u : 1u;
  }
}

template <int NV>
__global__ __launch_bounds__(BLOCK) void k_cell_compact(const uint32_t* cells, const uint32_t* keep,
                                                        const uint32_t* pos, uint64_t nc, uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nc || !keep[i]) return;
#pragma unroll
  for (int q = 0; q < NV; ++q) out[(uint64_t)pos[i] * NV + q] = cells[i * NV + q];
}

}  // namespace ftk2h
