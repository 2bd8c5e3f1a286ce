// Steps of a streaming extraction (ftk2h_stream_*), callable one by one so that the
// server can run the pending slabs of many streams as one batch.
#pragma once
#include "internal.hpp"

namespace ftk2h_internal {

// Smallest capacity of each of a stream's five committed-payload device arrays
// (keys, lambda, union-find parents, cells, roots).
constexpr size_t kDevVecMinBytes = (size_t)64 << 10;

// Copy the next timestep into its device half.  Fails if a slab is pending.
int stream_put(ftk2h_stream* S, const ftk2h_field_view* comps);
// A slab [t-2, t-1] awaits extraction.
bool stream_slab_ready(const ftk2h_stream* S);
// Mesh and device views of the pending slab (views[ncomp]).
void stream_slab_desc(const ftk2h_stream* S, ftk2h_mesh_desc* md, ftk2h_field_view* views);
// Hand the pending slab's extraction to the stream (takes ownership of R).
int stream_take_slab(ftk2h_stream* S, ftk2h_result* R);
// The same with the count of R's nodes on the slab's lower time plane already
// known (stream_launch_lower_count, for many slabs with one copy back).
int stream_take_slab_n(ftk2h_stream* S, ftk2h_result* R, uint32_t nlow);
void stream_launch_lower_count(const ftk2h_stream* S, const ftk2h_result* R, uint32_t* out, hipStream_t st);
uint64_t stream_timesteps(const ftk2h_stream* S);
// The HIP stream a stream's timestep copies and re-indexing run on.
hipStream_t stream_hip(const ftk2h_stream* S);
// Device memory held by the stream (timestep halves + committed nodes/cells).
uint64_t stream_device_bytes(const ftk2h_stream* S);

}  // namespace ftk2h_internal
