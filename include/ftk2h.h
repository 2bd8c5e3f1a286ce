/*
 * ftk2h.h — C ABI of the MI355X (gfx950) simplicial extraction library.
 *
 * This is the device boundary behind FTK2's SimplicialEngine<T, Predicate>:
 * include/ftk2/core/engine.hpp in this repository implements execute() and
 * get_complex() on top of these entry points.  Plain C types only: pointers,
 * sizes, status codes.  Every entry point returns an ftk2h_status; on failure
 * ftk2h_last_error() holds a message for the calling thread.  Nothing here calls
 * exit(); nothing here falls back to a CPU implementation.
 *
 * Reference interfaces replaced (hguo/ftk2, paths relative to its root):
 *   ftk2h_extract_regular      SimplicialEngine::execute()
 *                              include/ftk2/core/engine.hpp:96 (feed :101-212) and
 *                              SimplicialEngine::execute_cuda_regular :302-425
 *   ftk2h_result_*             SimplicialEngine::get_complex() engine.hpp:822-996,
 *                              FeatureElement feature.hpp:22-39,
 *                              FeatureComplex + sort() complex.hpp:14-58
 *   ftk2h_predicate_desc       CriticalPointPredicate<M,T> predicate.hpp:49-130,
 *                              ContourPredicate<T> :135-183, FiberPredicate<T> :340-389
 *                              (fields sos_q :25, threshold(s), scalar_var_name)
 *   ftk2h_mesh_desc            RegularSimplicialMesh(dims, offset, global_dims)
 *                              include/ftk2/core/mesh.hpp:256-265
 *   ftk2h_field_view           ftk::ndarray<T> read through get_value
 *                              engine.hpp:999-1009 (hguo/ndarray, un-vendored)
 *   ftk2h_stream_*             SimplicialEngine::execute_stream engine.hpp:214-270 and
 *                              execute_cuda_streaming_regular :601-819
 *   ftk2h_batch_*, ftk2h_server_*  no counterpart: batching of many meshes and a
 *                              request server over streams (each result equals the
 *                              single-mesh / single-stream entry point's)
 */
#ifndef FTK2H_H
#define FTK2H_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FTK2H_OK = 0,
  FTK2H_E_INVALID = 1,     /* bad argument (null pointer, bad dims, bad kind) */
  FTK2H_E_UNSUPPORTED = 2, /* (predicate, ndims, codimension, dtype) not built */
  FTK2H_E_OVERFLOW = 3,    /* a fixed capacity in ftk2h_options was too small */
  FTK2H_E_OOM = 4,         /* device allocation failed */
  FTK2H_E_HIP = 5,         /* a HIP runtime call failed */
  FTK2H_E_AGAIN = 6        /* not now: device-memory budget in use or result not ready */
} ftk2h_status;

typedef enum {
  FTK2H_PRED_CRITICAL_POINT = 0, /* CriticalPointPredicate<M,T>, codimension M = 1..3 */
  FTK2H_PRED_CONTOUR = 1,        /* ContourPredicate<T>, codimension 1 */
  FTK2H_PRED_FIBER = 2           /* FiberPredicate<T>, codimension 2 */
} ftk2h_pred_kind;

typedef enum { FTK2H_F32 = 0, FTK2H_F64 = 1 } ftk2h_dtype;

typedef struct ftk2h_ctx ftk2h_ctx;
typedef struct ftk2h_result ftk2h_result;

/* RegularSimplicialMesh: the last axis is time.  Vertex id =
 * x + gx*(y + gy*(z + gz*t)) over global dims of (local + offset) coordinates. */
typedef struct {
  int ndims;               /* total dimension d = 2..4 */
  uint64_t local_dims[4];  /* vertices per axis of this mesh (slab) */
  uint64_t offset[4];      /* global coordinate of local (0,...,0) */
  uint64_t global_dims[4]; /* extents used for vertex ids */
  uint64_t time_id_stride; /* 0: vertex id = x + gx*(y + gy*(z + gz*t)) (RegularSimplicialMesh).
                              Nonzero: the last axis is a layer of an ExtrudedSimplicialMesh
                              over a regular base, id = spatial id + t * time_id_stride with
                              time_id_stride = n_spatial_verts_ (mesh.hpp:544-560); must be at
                              least the spatial vertex count.  ftk2h_extract_regular only. */
} ftk2h_mesh_desc;

typedef struct {
  int kind;              /* ftk2h_pred_kind */
  int codimension;       /* m */
  int dtype;             /* ftk2h_dtype of every field view */
  double sos_q;          /* Predicate::sos_q, default 1e6 */
  double thresholds[2];  /* contour: [0]; fiber: [0],[1]; CP: unused */
} ftk2h_predicate_desc;

/* One scalar component indexed by LOCAL vertex coordinates:
 * element(l) = ptr[sum_j l_j * strides[j]]  (strides in elements).
 * The reference's legacy separate-array mode reads exactly this. */
typedef struct {
  const void* ptr;
  int64_t strides[4];
  int on_device; /* 1: device pointer read in place; 0: host memory, copied in */
} ftk2h_field_view;

typedef struct {
  uint64_t node_capacity; /* 0: size from the data; else fail with E_OVERFLOW beyond */
  uint64_t cell_capacity; /* same for connectivity cells */
  int sync;               /* 1: wait for completion before returning (default behaviour
                             when the struct is NULL) */
} ftk2h_options;

/* Device-resident view of a result (pointers stay owned by the result and valid
 * until ftk2h_result_free; device work that reads them must be complete by then,
 * since the library may reuse the memory for its next extraction). */
typedef struct {
  uint64_t n_nodes;
  int node_dim;              /* m */
  const uint64_t* keys;      /* canonical node keys: v_min * n_types + type, sorted */
  const void* lambda;        /* n_nodes * (m+1) values of the predicate dtype */
  const uint32_t* track_id;  /* minimum canonical index of the node's component */
  const float* scalar;       /* per node, or NULL */
  int n_conn;
  int conn_dim[3];
  uint64_t conn_cells[3];
  const uint32_t* conn[3];   /* conn_cells * (conn_dim+1) canonical indices, sorted */
  float scalar_value;        /* FeatureElement::scalar of every node when scalar is NULL
                                (the contour threshold, else 0) */
  int n_types;               /* key = v_min * n_types + type */
  uint64_t global_dims[4];   /* vertex id extents ... */
  uint64_t time_id_stride;   /* ... and the last axis' id step when extruded (else 0) */
} ftk2h_device_view;

int ftk2h_ctx_create(int device, ftk2h_ctx** out);
void ftk2h_ctx_destroy(ftk2h_ctx* ctx);

/* execute() + get_complex() on a RegularSimplicialMesh.  comps: ncomp views
 * (CP: M components, contour: 1, fiber: 2).  scalar: optional CP scalar field
 * interpolated at the feature (predicate.hpp:81-92), indexed like the reference
 * (global coordinates: the caller passes a view whose local index equals the
 * global one).  hip_stream: a hipStream_t or NULL for the context stream. */
int ftk2h_extract_regular(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh,
                          const ftk2h_predicate_desc* pred, const ftk2h_field_view* comps,
                          int ncomp, const ftk2h_field_view* scalar,
                          const ftk2h_options* opt, ftk2h_result** out, void* hip_stream);

int ftk2h_result_sizes(const ftk2h_result* r, uint64_t* n_nodes, int* node_dim, int* n_conn);
int ftk2h_result_conn_info(const ftk2h_result* r, int i, int* dim, uint64_t* n_cells);
/* Host copies; any pointer may be NULL.  vertices: n*(m+1) sorted vertex ids (the
 * Simplex of each FeatureElement); lambda: n*(m+1) doubles (solver output before
 * the float cast); bary: n*(m+1) floats (FeatureElement::barycentric_coords[0]);
 * scalar: n floats; track_id: n (FeatureElement::track_id). */
int ftk2h_result_copy_nodes(const ftk2h_result* r, uint64_t* vertices, double* lambda,
                            float* bary, float* scalar, uint64_t* track_id);
/* The same for nodes [begin, begin+count) only (E_INVALID past the end): reads a
 * sample of a large result without copying all of it.  No counterpart in the
 * reference, whose get_complex() always copies every FeatureElement. */
int ftk2h_result_copy_node_range(const ftk2h_result* r, uint64_t begin, uint64_t count,
                                 uint64_t* vertices, double* lambda, float* bary, float* scalar,
                                 uint64_t* track_id);
int ftk2h_result_copy_keys(const ftk2h_result* r, uint64_t* keys);
int ftk2h_result_copy_conn(const ftk2h_result* r, int i, uint32_t* indices);
int ftk2h_result_device_view(const ftk2h_result* r, ftk2h_device_view* out);
/* Per-phase device time in ms of the extraction that produced r (n entries, HIP events
 * around each kernel on the extraction's stream; phases: classification stage 1,
 * stage 2, node scan, offsets + cell count, cell scan + host sync, emit, link,
 * compress). */
int ftk2h_result_phase_ms(const ftk2h_result* r, float* ms, int n);
void ftk2h_result_free(ftk2h_result* r);

/* ---- Attributes (engine.hpp:1144-1201, AttributeSpec predicate.hpp:34-41) --------
 * FeatureElement::attributes[slot] = (float) sum_i bary_i * value_i over the node's
 * vertices in ascending id order, bary_i the float barycentric coordinates, the sum
 * in the predicate dtype.  value_i = comps[0] at the vertex (ncomp == 1) or the
 * Euclidean norm of comps[0..ncomp) (ncomp 2..16, the reference's "magnitude").
 * Views are indexed by local vertex coordinates of the extraction's mesh; the
 * reference's choice of array, component and coordinates is the caller's (the C++
 * engine in include/ftk2/core/engine.hpp makes it as the reference does).  Later
 * entries overwrite earlier ones with the same slot. */
typedef struct {
  int slot;                      /* 0..15 */
  int ncomp;                     /* 1..16 */
  ftk2h_field_view comps[16];
} ftk2h_attribute_desc;
int ftk2h_result_add_attributes(ftk2h_ctx* ctx, ftk2h_result* r, const ftk2h_mesh_desc* mesh,
                                const ftk2h_attribute_desc* attrs, int nattr, void* hip_stream);
/* n_nodes * 16 floats (slots without an attribute are 0). */
int ftk2h_result_copy_attributes(const ftk2h_result* r, float* out);

/* Feature positions as the reference's VTK writer computes them
 * (utils/vtk.hpp:26-59): xyz[3*i + j] = (float) sum_k (double)bary_k * coord_j of
 * the node's vertices (global grid indices), first three coordinates, 0 beyond
 * the mesh dimension; time[i] (may be NULL) = the last coordinate when the mesh is
 * extruded (time_id_stride != 0), else 0.  Host buffers of 3 * n_nodes and
 * n_nodes floats. */
int ftk2h_result_copy_positions(const ftk2h_result* r, float* xyz, float* time);

/* ---- Critical-point classification (opt-in) ---------------------------------------
 * No counterpart in the reference's engine, whose CriticalPointPredicate always sets
 * FeatureElement::type = 0 (predicate.hpp:80); its tracking configuration describes
 * Hessian classification of gradient fields (min / max / saddles) and Jacobian
 * classification of flow fields (source / sink / spiral), high_level/
 * tracking_config.hpp:84-88.  For critical points of an M-component field on an
 * M-dimensional space + time (M = 1..3): the spatial Jacobian J of the field at each
 * vertex of the node's simplex by central differences over the extraction's local
 * grid (one-sided at its faces, per timestep), combined as sum_k lambda_k J_k with the
 * node's barycentric coordinates, in double.  GRADIENT classifies the symmetric part
 * (J + J^T)/2 by the signs of its eigenvalues; FLOW classifies J's eigenvalues
 * (real parts, and a complex pair as a spiral).  comps: the views of the extraction.
 * Types stay 0 unless this is called; ftk2h_result_copy_types reads them. */
typedef enum { FTK2H_CLASSIFY_GRADIENT = 1, FTK2H_CLASSIFY_FLOW = 2 } ftk2h_classify_mode;
typedef enum {
  FTK2H_CP_UNCLASSIFIED = 0,
  FTK2H_CP_MINIMUM = 1,       /* gradient: every eigenvalue positive */
  FTK2H_CP_MAXIMUM = 2,       /* gradient: every eigenvalue negative */
  FTK2H_CP_SADDLE = 3,        /* gradient, M = 2: one of each sign */
  FTK2H_CP_SADDLE_1 = 4,      /* gradient, M = 3: one negative eigenvalue */
  FTK2H_CP_SADDLE_2 = 5,      /* gradient, M = 3: two negative eigenvalues */
  FTK2H_CP_SOURCE = 16,       /* flow: every real part positive, real eigenvalues */
  FTK2H_CP_SINK = 17,         /* flow: every real part negative, real eigenvalues */
  FTK2H_CP_FLOW_SADDLE = 18,  /* flow: real parts of both signs */
  FTK2H_CP_SPIRAL_SOURCE = 19,/* flow: complex pair, every real part positive */
  FTK2H_CP_SPIRAL_SINK = 20,  /* flow: complex pair, every real part negative */
  FTK2H_CP_CENTER = 21,       /* flow: a purely imaginary pair */
  FTK2H_CP_DEGENERATE = 255   /* a zero eigenvalue (det J = 0) */
} ftk2h_cp_type;
int ftk2h_result_classify(ftk2h_ctx* ctx, ftk2h_result* r, const ftk2h_mesh_desc* mesh,
                          const ftk2h_field_view* comps, int ncomp, int mode, void* hip_stream);
/* n_nodes types (FeatureElement::type); zeros when not classified. */
int ftk2h_result_copy_types(const ftk2h_result* r, uint32_t* types);

/* ---- Streaming (SimplicialEngine::execute_stream, engine.hpp:214-270) ----------
 * Timesteps of a spatial grid arrive one at a time (ftk2h_stream_append, views
 * indexed by spatial coordinates).  Each append of timestep t >= 1 extracts the
 * slab [t-1, t] on the device; the last timestep and the open union-find frontier
 * stay resident.  ftk2h_stream_finish returns the complex over all timesteps, the
 * same as one ftk2h_extract_regular over the whole spacetime grid. */
typedef struct ftk2h_stream ftk2h_stream;
int ftk2h_stream_open(ftk2h_ctx* ctx, int spatial_ndims, const uint64_t* spatial_dims,
                      const ftk2h_predicate_desc* pred, int ncomp, ftk2h_stream** out);
int ftk2h_stream_append(ftk2h_stream* s, const ftk2h_field_view* comps, void* hip_stream);
int ftk2h_stream_finish(ftk2h_stream* s, ftk2h_result** out);
void ftk2h_stream_close(ftk2h_stream* s);
/* Retirement of finished stream state.  No counterpart in the reference, whose
 * streaming path copies every slab's result to host memory and links it there
 * (engine.hpp:756-800).  When the committed nodes, lambda and cells held on the
 * device exceed `bytes` after a slab, they move to host memory and the device keeps
 * only the open frontier: parents of the last slab's lower-plane nodes and of the
 * roots of the tracks that still reach it.  ftk2h_stream_finish returns the same
 * result either way.  Default 8 GiB; 0 retires after every slab; UINT64_MAX never.
 * Retirement copies through pageable host memory and waits for the copy, so a low
 * threshold trades append time for device memory. */
int ftk2h_stream_set_retire_bytes(ftk2h_stream* s, uint64_t bytes);
/* Windows of several slabs.  No counterpart in the reference, whose streaming path
 * extracts one slab [t-1, t] per timestep (engine.hpp:214-270).  With `timesteps` =
 * W + 1 > 2 the stream keeps W + 1 timesteps on the device and extracts a window of
 * W slabs once it is full (and the last, partial one in ftk2h_stream_finish): one
 * mesh of local time extent W + 1, so every plane inside a window is classified
 * once instead of twice.  The result is the same.  Set before the first append;
 * default 2 (one slab per timestep, as the reference). */
int ftk2h_stream_set_window(ftk2h_stream* s, int timesteps);
/* Device bytes in use by the stream: its window's timestep buffers and the committed
 * nodes, lambda, cells and union-find parents still on the device (allocations
 * grow to at most about the retirement threshold plus one slab). */
uint64_t ftk2h_stream_resident_bytes(const ftk2h_stream* s);
/* Committed nodes moved to host memory so far. */
uint64_t ftk2h_stream_retired_nodes(const ftk2h_stream* s);

/* ---- Batching -------------------------------------------------------------------
 * No counterpart in the reference, which runs one SimplicialEngine::execute()
 * (engine.hpp:96) per mesh.  Many independent RegularSimplicialMesh jobs of one
 * predicate (same kind, codimension, dtype and mesh dimension) are extracted in one
 * pass: one job table on the device, the jobs' tiles scheduled over one persistent
 * grid, the launch sequence captured in a hipGraph and replayed while its shape is
 * unchanged.  results[j] equals ftk2h_extract_regular on job j alone, bit for bit.
 * The results of one run share device memory; each is freed by ftk2h_result_free.
 * Predicates whose cells are star-fan triangles (codimension d - 2 with d - m = 2)
 * are not batched (FTK2H_E_UNSUPPORTED). */
typedef struct ftk2h_batch ftk2h_batch;
typedef struct {
  ftk2h_mesh_desc mesh;
  const ftk2h_field_view* comps;  /* as many views as the predicate reads */
  const ftk2h_field_view* scalar; /* optional CP scalar field, or NULL */
} ftk2h_job;
typedef struct {
  int grid_per_cu; /* persistent workgroups per compute unit; 0: one workgroup per tile */
  int chunk;       /* with a persistent grid: 0 = static stride, else tiles per atomic grab */
  int use_graph;   /* 1: capture the launches in a hipGraph and replay it */
} ftk2h_batch_config;
int ftk2h_batch_create(ftk2h_ctx* ctx, const ftk2h_predicate_desc* pred, int ndims, ftk2h_batch** out);
int ftk2h_batch_set_config(ftk2h_batch* b, const ftk2h_batch_config* cfg);
int ftk2h_batch_get_config(const ftk2h_batch* b, ftk2h_batch_config* cfg);
/* hip_stream: NULL for the batch's own stream.  Returns when every result is ready. */
int ftk2h_batch_run(ftk2h_batch* b, int njobs, const ftk2h_job* jobs, ftk2h_result** results, void* hip_stream);
/* Times every candidate schedule on these jobs (reps runs each, best of), keeps the
 * fastest as the batch's configuration and returns it with its time in *ms. */
int ftk2h_batch_autotune(ftk2h_batch* b, int njobs, const ftk2h_job* jobs, int reps, ftk2h_batch_config* best,
                         float* ms);
/* Device ms of the last run: [0] classify..job bases, [1] emit..link, [2] host wall. */
int ftk2h_batch_last_ms(const ftk2h_batch* b, float* ms, int n);
void ftk2h_batch_destroy(ftk2h_batch* b);

/* ---- Serving ----------------------------------------------------------------------
 * No counterpart in the reference.  Requests are streams (execute_stream semantics,
 * engine.hpp:214-270) admitted against a device-memory budget.  Each
 * ftk2h_server_step extracts the next slab of every admitted stream that has a
 * queued timestep, the slabs of one predicate template as one ftk2h_batch_run
 * (continuous batching: streams join at the first step after admission and leave
 * when closed and drained).  A stream's result equals the one it gets alone. */
typedef struct ftk2h_server ftk2h_server;
int ftk2h_server_create(ftk2h_ctx* ctx, uint64_t hbm_budget_bytes, ftk2h_server** out);
/* Admit a stream of at most max_timesteps timesteps.  FTK2H_E_AGAIN: its device
 * reservation does not fit the budget now (retry after results are taken);
 * FTK2H_E_OVERFLOW: it can never fit. */
int ftk2h_server_open(ftk2h_server* s, int spatial_ndims, const uint64_t* spatial_dims,
                      const ftk2h_predicate_desc* pred, int ncomp, uint64_t max_timesteps, uint64_t* id);
/* Queue one timestep (views indexed by spatial coordinates).  Host views are copied
 * at once; device views are read by the step that consumes them. */
int ftk2h_server_append(ftk2h_server* s, uint64_t id, const ftk2h_field_view* comps);
int ftk2h_server_close(ftk2h_server* s, uint64_t id); /* no more timesteps */
/* One continuous-batching step; done[0..min(n_done, max_done)) = streams finished in it. */
int ftk2h_server_step(ftk2h_server* s, uint64_t* done, int max_done, int* n_done);
/* Result of a finished stream (FTK2H_E_AGAIN before); releases its reservation. */
int ftk2h_server_take(ftk2h_server* s, uint64_t id, ftk2h_result** out);
int ftk2h_server_stats(const ftk2h_server* s, uint64_t* reserved_bytes, uint64_t* device_bytes, int* active,
                       int* queued_timesteps);
void ftk2h_server_destroy(ftk2h_server* s);

/* ---- Extraction without device allocation ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same extraction as ftk2h_extract_regular, with the result written
 * into the caller's device arrays and every work array (per-vertex masks and offsets,
 * tile counts, union-find parents, cross-workgroup edges) taken from one fixed
 * device array of FTK2H_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from
 * the first ftk2h_ctx_create on that GPU until the process ends.  Calls no hipMalloc
 * or hipFree.  Field views must be device views (on_device = 1).  Not built for
 * star-fan cells (codimension 2 on 3-D + time meshes: FTK2H_E_UNSUPPORTED).
 * Every call on one GPU uses that GPU's scratch array: calls that overlap on one GPU
 * (other host threads, other contexts or streams) race and must be serialized by
 * the caller.  FTK2H_E_OVERFLOW when the mesh's work arrays exceed the scratch array
 * or the counts exceed the capacities below (*n_nodes and *n_cells then hold the
 * required counts when they are known). */
#define FTK2H_NOALLOC_SCRATCH_BYTES (384ull << 20)
typedef struct {
  uint64_t* keys;         /* node_capacity node keys (v_min * n_types + type), sorted */
  void* lambda;           /* node_capacity * (m+1) values of the predicate dtype */
  uint32_t* track_id;     /* node_capacity: minimum canonical index of the component */
  float* scalar;          /* node_capacity, required when a scalar field is given, else NULL */
  uint32_t* cells;        /* cell_capacity * (cell dim + 1) canonical indices, sorted */
  uint64_t node_capacity;
  uint64_t cell_capacity;
} ftk2h_output_buffers;
int ftk2h_extract_regular_noalloc(ftk2h_ctx* ctx, const ftk2h_mesh_desc* mesh, const ftk2h_predicate_desc* pred,
                                  const ftk2h_field_view* comps, int ncomp, const ftk2h_field_view* scalar,
                                  const ftk2h_output_buffers* out, uint64_t* n_nodes, uint64_t* n_cells,
                                  void* hip_stream);

const char* ftk2h_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FTK2H_H */
