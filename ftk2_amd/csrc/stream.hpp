// Steps of a streaming extraction (ftk2h_stream_*), callable one by one so that the
// server can run the pending slabs of many streams as one batch.
#pragma once
#include "internal.hpp"

namespace ftk2h_internal {

// Smallest capacity of each of a stream's five committed-payload device arrays
// (keys, lambda, union-find parents, cells, roots).
constexpr size_t kDevVecMinBytes = (size_t)64 << 10;

// Copy the next timestep into its window slot.  Fails if the window is full.
int stream_put(ftk2h_stream* S, const ftk2h_field_view* comps);
// The window is full (it holds as many timesteps as its slots): it must be
// extracted before the next timestep.
bool stream_slab_ready(const ftk2h_stream* S);
// The window holds at least two timesteps (one slab or more) to extract.
bool stream_window_pending(const ftk2h_stream* S);
// Timesteps of a full window (>= 2; before the first timestep only).
int stream_set_window(ftk2h_stream* S, int planes);
// Mesh and device views of the pending window (views[ncomp]).
void stream_slab_desc(const ftk2h_stream* S, ftk2h_mesh_desc* md, ftk2h_field_view* views);
// Hand the pending window's extraction to the stream (takes ownership of R).
int stream_take_slab(ftk2h_stream* S, ftk2h_result* R);
// The same with the count of R's nodes below the window's last time plane already
// known (stream_launch_lower_count, for many windows with one copy back).
int stream_take_slab_n(ftk2h_stream* S, ftk2h_result* R, uint32_t nlow);
void stream_launch_lower_count(const ftk2h_stream* S, const ftk2h_result* R, uint32_t* out, hipStream_t st);
// Node keys below this one have their smallest vertex before the window's last plane.
uint64_t stream_lower_bound_key(const ftk2h_stream* S);
uint64_t stream_timesteps(const ftk2h_stream* S);
// The HIP stream a stream's timestep copies and re-indexing run on.
hipStream_t stream_hip(const ftk2h_stream* S);
// Device memory held by the stream (window slots + committed nodes/cells).
uint64_t stream_device_bytes(const ftk2h_stream* S);

}  // namespace ftk2h_internal
