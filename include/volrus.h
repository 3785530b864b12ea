/* volrus.h — C ABI of the MI355X (gfx950) narrow-band level-set library.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes.  Every call
 * returns VR_OK (0) or an error code and never aborts; the text of the last
 * error on a context is available from vr_last_error().
 *
 * Grids cross the boundary in the reference's NanoGrid byte layout
 * (src/nano/layout.rs:13-52: 320-byte header, then 2128-byte leaves sorted
 * by origin), so a Rust shim can pass NanoGrid::as_bytes() straight in and
 * rebuild a Grid with NanoGrid::from_bytes(..).to_grid() on the way out.
 *
 * Threading: one context owns one HIP stream; calls on a context must be
 * serialised by the caller, different contexts may be used concurrently.
 * Kernels are asynchronous; calls that return sizes or copy to host memory
 * synchronise the context's stream.
 */
#ifndef VOLRUS_H
#define VOLRUS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VR_OK = 0,
  VR_ERR_INVALID_ARGUMENT = 1,
  VR_ERR_INVALID_NANO = 2, /* NanoGrid::from_bytes would return Err */
  VR_ERR_OUT_OF_MEMORY = 3,
  VR_ERR_HIP = 4,
  VR_ERR_UNSUPPORTED = 5,
  VR_ERR_CAPACITY = 6, /* caller buffer too small; required size reported */
  VR_ERR_INVALID_VOL = 7 /* read_vol would return Err (InvalidData / UnexpectedEof) */
};

/* Integration schemes for vr_advect. */
enum { VR_SCHEME_RK1 = 1, VR_SCHEME_RK3 = 3 };

/* Analytic velocity kinds for vr_velocity_analytic. */
enum {
  VR_VEL_UNIFORM = 2, /* u = (p0, p1, p2) */
  VR_VEL_RIGID = 3,   /* u = w x (x - c), w = (p0, p1, p2), c = (p3, p4, p5) */
  VR_VEL_VORTEX = 4   /* u = (sin^2(pi x) sin(2 pi y), -sin^2(pi y) sin(2 pi x), 0) */
};

typedef struct vr_ctx vr_ctx;
typedef struct vr_grid vr_grid;
typedef struct vr_velocity vr_velocity;
typedef struct vr_mesh vr_mesh;

/* ---- context ------------------------------------------------------------ */
/* No counterpart in the reference (single-threaded CPU code). */
int vr_ctx_create(int device, vr_ctx** out);
int vr_ctx_destroy(vr_ctx* ctx);
const char* vr_last_error(const vr_ctx* ctx);
int vr_ctx_synchronize(vr_ctx* ctx);
/* The context's hipStream_t, for callers that time or order work on it. */
void* vr_ctx_stream(vr_ctx* ctx);

/* ---- grids (NanoGrid boundary) ------------------------------------------ */
/* Replaces NanoGrid::from_bytes (src/nano/layout.rs:202-251) plus the upload
 * the reference never performs: validates magic, version, size and leaf
 * order with the same rules, then copies the leaves to HBM and builds the
 * root/internal/neighbour lookup arrays.  Also rejects leaf origins that are
 * not multiples of 8 (NanoGrid::from_grid never writes them). */
int vr_grid_upload_nano(vr_ctx* ctx, const uint8_t* bytes, size_t len, vr_grid** out);
/* Replaces NanoGrid::from_grid + as_bytes (src/nano/layout.rs:89-170,197-199)
 * for a device grid: writes exactly the bytes from_grid would produce for the
 * grid's contents.  With out == NULL or cap too small, *len receives the size
 * and VR_ERR_CAPACITY is returned (NULL/0 is the size query, returns VR_OK). */
int vr_grid_download_nano(vr_ctx* ctx, const vr_grid* g, uint8_t* out, size_t cap, size_t* len);
/* read_vol (src/io/vol_format.rs:88-147) plus the upload: the same checks
 * and messages (bad magic, version, grid class, short buffer ->
 * VR_ERR_INVALID_VOL); each record becomes the leaf at origin & ~7, a later
 * record of a leaf replacing an earlier one; uniform transform, zero origin,
 * as Grid::new. */
int vr_grid_upload_vol(vr_ctx* ctx, const uint8_t* bytes, size_t len, vr_grid** out);
/* write_vol (src/io/vol_format.rs:53-85) for a device grid, leaves in
 * ascending origin order (the reference writes its HashMap order).  Size
 * query and VR_ERR_CAPACITY as vr_grid_download_nano. */
int vr_grid_download_vol(vr_ctx* ctx, const vr_grid* g, uint8_t* out, size_t cap, size_t* len);
/* Grid::leaf_count / active_voxel_count (src/grid.rs:505-513). */
int vr_grid_counts(vr_ctx* ctx, const vr_grid* g, uint64_t* leaf_count, uint64_t* active_count);
int vr_grid_destroy(vr_ctx* ctx, vr_grid* g);
/* Device-side deep copy of a grid (leaves and lookup arrays). */
int vr_grid_clone(vr_ctx* ctx, const vr_grid* g, vr_grid** out);

/* ---- synthetic level sets (fixture generation on the device) ----------- */
enum { VR_PRIM_SPHERE = 0, VR_PRIM_BOX = 1, VR_PRIM_NOISY_SPHERE = 2 };
enum { VR_CSG_UNION = 0, VR_CSG_INTERSECTION = 1, VR_CSG_DIFFERENCE = 2 };
/* sphere:       p = (cx, cy, cz, r)
 * box:          p = (xmin, ymin, zmin, xmax, ymax, zmax), exact box distance
 * noisy sphere: p = (cx, cy, cz, r, amplitude, base frequency, seed)
 * `op` combines primitive i (i >= 1) with the result of 0..i-1. */
typedef struct vr_primitive {
  int32_t kind;
  int32_t op;
  double p[8];
} vr_primitive;
/* Builds ((p0 op1 p1) op2 p2) ... on the device, each primitive rasterised
 * like make_level_set_sphere (src/tools/sphere.rs:17-59) and combined like
 * csg_union / csg_intersection / csg_difference (src/tools/csg.rs:15-78). */
int vr_grid_from_primitives(vr_ctx* ctx, const vr_primitive* prims, int count, double voxel_size, double half_width,
                            vr_grid** out);

/* ---- velocity ----------------------------------------------------------- */
/* VelocityField { vx, vy, vz } (src/tools/advection.rs:10-24): three uploaded
 * grids, each sampled with sample_trilinear.  The grids are borrowed and must
 * outlive the velocity object. */
int vr_velocity_grids3(vr_ctx* ctx, const vr_grid* vx, const vr_grid* vy, const vr_grid* vz, vr_velocity** out);
/* Dense n^3 field per component (3 * n^3 floats, component-major, index
 * (i * n + j) * n + k) at nodes (i h, j h, k h), background 0 outside; sampled
 * like sample_trilinear on a grid of voxel size h.  No reference counterpart. */
int vr_velocity_dense(vr_ctx* ctx, const float* uvw, int n, double h, vr_velocity** out);
/* Analytic fields evaluated in double on the device.  No reference counterpart. */
int vr_velocity_analytic(vr_ctx* ctx, int kind, const double params[6], vr_velocity** out);
/* Multiply the field by cospi(t / period) at each stage time (period > 0),
 * or by 1 (period <= 0, the default). */
int vr_velocity_set_period(vr_velocity* v, double period);
int vr_velocity_destroy(vr_ctx* ctx, vr_velocity* v);

/* ---- the path ----------------------------------------------------------- */
/* advect_level_set(&Grid<f32>, &VelocityField, dt) -> Grid<f32>
 * (src/tools/advection.rs:32-56) for scheme VR_SCHEME_RK1 and a time-free
 * field; VR_SCHEME_RK3 (SSP back-trace) and time factors are extensions.
 * t is the time at the start of the step. */
int vr_advect(vr_ctx* ctx, const vr_grid* in, const vr_velocity* vel, double dt, double t, int scheme,
              vr_grid** out);
/* find_zero_crossings(&Grid<f32>) -> Vec<Coord> (src/tools/level_set_track.rs:14-47).
 * Writes *count coordinates (x, y, z) in leaf order into xyz (3 * cap ints). */
int vr_zero_crossings(vr_ctx* ctx, const vr_grid* g, int32_t* xyz, size_t cap, size_t* count);
/* trim_narrow_band(&mut Grid<f32>, half_width) (src/tools/level_set_track.rs:53-68). */
int vr_trim_band(vr_ctx* ctx, vr_grid* g, float half_width);
/* rebuild_narrow_band(&mut Grid<f32>, half_width) (src/tools/level_set_track.rs:80-114).
 * Requires ceil(half_width) <= 8 (VR_ERR_UNSUPPORTED otherwise). */
int vr_rebuild_band(vr_ctx* ctx, vr_grid* g, float half_width);
/* extend_velocity(&sdf, &mut vx, &mut vy, &mut vz, half_width)
 * (src/tools/level_set_track.rs:127-169): every band voxel of sdf
 * (vs <= |phi| < half_width * vs, active) takes, in vx, vy and vz (activated,
 * with a new leaf where a grid has none), the values at the interface voxel
 * (|phi| < vs, active) at the least squared index distance; among equals the
 * first in iteration order (internal-node origin ascending, then leaf slot,
 * then voxel offset: the reference's order inside one 128^3 node).  vx, vy,
 * vz must be distinct grids. */
int vr_extend_velocity(vr_ctx* ctx, const vr_grid* sdf, vr_grid* vx, vr_grid* vy, vr_grid* vz, float half_width);
/* volume_to_mesh(&Grid<f32>, iso) -> Mesh (src/tools/volume_to_mesh.rs:46-101). */
int vr_marching_cubes(vr_ctx* ctx, const vr_grid* g, float iso, vr_mesh** out);
/* Mesh::vertex_count / tri_count (src/tools/volume_to_mesh.rs:18-25). */
int vr_mesh_counts(vr_ctx* ctx, const vr_mesh* m, uint64_t* vertex_count, uint64_t* tri_count);
/* Copies Mesh::vertices ([f64; 3] each) and Mesh::triangles ([u32; 3] each). */
int vr_mesh_copy(vr_ctx* ctx, const vr_mesh* m, double* xyz, size_t vcap, uint32_t* tri, size_t tcap);
int vr_mesh_destroy(vr_ctx* ctx, vr_mesh* m);
/* An empty mesh object for vr_rollout to fill (and reuse across calls). */
int vr_mesh_create(vr_ctx* ctx, vr_mesh** out);

/* No counterpart in the reference: volume_to_mesh for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  It calls no
 * hipMalloc or hipFree and synchronises nothing: its scratch is a fixed
 * array in device memory (about 30 MB per device, held until the process
 * ends) for grids of at most VR_NOALLOC_MAX_LEAVES leaves (VR_ERR_CAPACITY
 * above that).  The caller passes device buffers: xyz (3 * vcap doubles),
 * tri (3 * tcap uint32) and counts (2 int32: vertices, triangles, written on
 * the context's stream).  Vertices and triangles past vcap / tcap are not
 * written, the counts are always the full ones; the output equals
 * vr_marching_cubes.  Calls that overlap on one device (two contexts on
 * different streams) share the array and race. */
#define VR_NOALLOC_MAX_LEAVES 4096
int vr_marching_cubes_noalloc(vr_ctx* ctx, const vr_grid* g, float iso, double* xyz, uint64_t vcap, uint32_t* tri,
                              uint64_t tcap, int32_t* counts);

/* ---- adaptive meshing ---------------------------------------------------- */
/* adaptive_surface(&Grid<f32>, &AdaptiveMeshConfig) -> RenderMesh
 * (src/tools/adaptive_mesh.rs:100-241): marching cubes, then max_depth rounds
 * of 1-to-4 midpoint subdivision of the triangles whose largest vertex |mean
 * curvature| exceeds curvature_threshold and whose longest edge exceeds
 * max_edge_length / 2^(depth+1), new vertices projected onto the isosurface;
 * angle-weighted vertex normals and signed curvatures.  Vertices, triangles
 * and curvatures follow the reference's arithmetic and numbering; normals
 * use the device's acos. */
typedef struct vr_render_mesh vr_render_mesh;
int vr_adaptive_surface(vr_ctx* ctx, const vr_grid* g, float isovalue, double max_edge_length,
                        double curvature_threshold, uint32_t max_depth, vr_render_mesh** out);
int vr_render_mesh_counts(vr_ctx* ctx, const vr_render_mesh* m, uint64_t* vertex_count, uint64_t* tri_count);
/* RenderMesh::{vertices, normals, curvatures} (vcap entries each, any may be
 * NULL) and triangles (tcap). */
int vr_render_mesh_copy(vr_ctx* ctx, const vr_render_mesh* m, double* xyz, double* normals, double* curvatures,
                        size_t vcap, uint32_t* tri, size_t tcap);
int vr_render_mesh_destroy(vr_ctx* ctx, vr_render_mesh* m);

/* ---- rollout ------------------------------------------------------------ */
/* No counterpart in the reference (nothing there chains its tools).  Runs
 * `steps` advection steps from time t0 (step s starts at t0 + s * dt), calls
 * rebuild_narrow_band(half_width) after every `rebuild_every`-th step and
 * volume_to_mesh(iso) after every `mesh_every`-th step (0 disables either),
 * in place on g.  The last mesh is left in `mesh` (may be NULL).  The call
 * synchronises with the host once, at the end; meshes on the way are written
 * into the mesh's buffers up to their capacity, and when a mesh is larger
 * the buffers grow and the state of the last mesh step is meshed again (at
 * that step, with one extra synchronisation, when it is not the last step),
 * so the returned mesh is always whole and always shows the state after step
 * (steps / mesh_every) * mesh_every (pass the same mesh to later calls to
 * keep the capacity). */
typedef struct vr_rollout_stats {
  uint64_t voxel_updates; /* sum over steps of the active voxels advected */
  uint64_t advect_steps;
  uint64_t rebuilds;
  uint64_t meshes;
  uint64_t last_vertex_count;
  uint64_t last_triangle_count;
  uint64_t final_leaf_count;
  uint64_t final_active_count;
} vr_rollout_stats;
int vr_rollout(vr_ctx* ctx, vr_grid* g, const vr_velocity* vel, double t0, double dt, int scheme, int steps,
               int rebuild_every, float half_width, int mesh_every, float iso, vr_mesh* mesh,
               vr_rollout_stats* stats);

/* ---- differentiable rollout -------------------------------------------- */
/* No counterpart in the reference (it has no derivatives).  `steps` advection
 * steps (scheme RK1 or RK3) of phi0 through a static dense velocity field,
 * topology fixed, then loss = 1/2 sum over active voxels of
 * (phi_steps - target.get(x))^2 and its gradient with respect to the dense
 * field, written to grad (3 n^3 floats, the field's layout; may be NULL).
 * Forward states are kept every `checkpoint_every` steps and the others are
 * recomputed in the reverse pass.  Each reverse step sums, per active voxel,
 * lambda times the gradient of its departure-point sample (12 bytes per
 * active voxel of device memory); one launch after the last reverse step
 * pulls the sums back onto the field's nodes.  The gradient is summed with
 * float atomics, so its last bits can vary from run to run. */
typedef struct vr_adjoint_stats {
  double loss;
  double forward_ms; /* device time of the forward steps */
  double adjoint_ms; /* device time of recomputation + reverse steps */
  uint64_t steps;
  uint64_t active_voxels;
} vr_adjoint_stats;
int vr_adjoint_rollout(vr_ctx* ctx, const vr_grid* phi0, const vr_velocity* vel, double dt, int scheme, int steps,
                       int checkpoint_every, const vr_grid* target, vr_adjoint_stats* stats, float* grad);
/* vr_adjoint_rollout for `count` level sets at once (each with its own dense
 * field, time step and target): every forward step, every backward step and
 * the final pull-back onto the fields is one launch over the leaves of all
 * grids.  stats[g] (count entries, may
 * be NULL) holds grid g's loss and the whole batch's forward / adjoint
 * times; grads[g] (may be NULL, or hold NULL entries) receives grid g's
 * gradient. */
int vr_adjoint_rollout_batch(vr_ctx* ctx, int count, const vr_grid* const* phi0, const vr_velocity* const* vels,
                             const double* dts, int scheme, int steps, int checkpoint_every,
                             const vr_grid* const* targets, vr_adjoint_stats* stats, float* const* grads);

/* ---- batches of level sets ----------------------------------------------- */
/* No counterpart in the reference.  A batch advects every grid with its own
 * velocity and meshes the results in one step: one launch over a global leaf
 * worklist cut into chunks of `leaves_per_grab` leaves of one grid and
 * partitioned over the 8 XCDs -- persistent, `waves_per_cu` waves per CU
 * taking chunks from per-XCD atomic counters, or, with waves_per_cu = 0, one
 * block per chunk -- followed by marching cubes over the concatenated tiles
 * of all grids.  The advect code is specialised for the velocity kind when
 * all grids share one.  The batch copies the input leaves at creation, so
 * the input grids may be changed or destroyed afterwards; the velocities are
 * borrowed and must outlive the batch.  Outputs live in the batch
 * (vr_batch_output).  A GRAPH step replays the graph captured for the same
 * dt, t, scheme, iso and tuning; anything that changes a buffer it uses
 * (vr_batch_set_dt, a new tuning, a larger mesh) captures it again.  Modes:
 *   VR_BATCH_LOOP  the per-grid path (one advect launch + one MC pipeline per grid),
 *   VR_BATCH_EAGER the batched kernels launched directly,
 *   VR_BATCH_GRAPH the batched kernels captured once into a hipGraph and replayed. */
typedef struct vr_batch vr_batch;
enum { VR_BATCH_LOOP = 0, VR_BATCH_EAGER = 1, VR_BATCH_GRAPH = 2 };
int vr_batch_create(vr_ctx* ctx, vr_grid* const* grids, const vr_velocity* const* vels, int count, vr_batch** out);
int vr_batch_set_tuning(vr_batch* b, int leaves_per_grab, int waves_per_cu);
/* Per-grid time steps (count doubles); NULL returns to the dt of vr_batch_step. */
int vr_batch_set_dt(vr_ctx* ctx, vr_batch* b, const double* dts);
/* Times the batched advect launch for leaves-per-grab in {1,2,4,8,16} and
 * waves-per-CU in {0 (one block per chunk),4,8,16,32} and keeps the fastest
 * pair. */
int vr_batch_autotune(vr_ctx* ctx, vr_batch* b, double dt, int scheme, int* leaves_per_grab, int* waves_per_cu);
int vr_batch_step(vr_ctx* ctx, vr_batch* b, double dt, double t, int scheme, float iso, int mode);
/* Mesh of grid i from the last step in `mode` (LOOP keeps its own meshes).
 * VR_ERR_CAPACITY if the mesh does not lie inside the batch's buffers. */
int vr_batch_mesh_counts(vr_ctx* ctx, vr_batch* b, int i, int mode, uint64_t* vertex_count, uint64_t* tri_count);
int vr_batch_mesh_copy(vr_ctx* ctx, vr_batch* b, int i, int mode, double* xyz, size_t vcap, uint32_t* tri,
                       size_t tcap);
int vr_batch_output(vr_batch* b, int i, const vr_grid** out);
int vr_batch_destroy(vr_ctx* ctx, vr_batch* b);

/* ---- measurement -------------------------------------------------------- */
/* Kernel classes for vr_ctx_timing (GRAPH: whole batch graph launches; MC:
 * the marching-cubes count pass, MC_EMIT: its vertex and triangle pass). */
enum {
  VR_TIMING_ADVECT = 0,
  VR_TIMING_CROSSINGS = 1,
  VR_TIMING_REBUILD = 2,
  VR_TIMING_COMPACT = 3,
  VR_TIMING_MC = 4,
  VR_TIMING_GRAPH = 5,
  VR_TIMING_MC_EMIT = 6
};
/* Turns HIP-event timing on the context stream on or off and clears the
 * totals.  Each launch is bracketed by its own pair of events, except the
 * back-to-back advect launches of a rollout segment (between two rebuild or
 * mesh passes), which one pair brackets as a run (launches counted singly). */
int vr_ctx_set_timing(vr_ctx* ctx, int enable);
/* Device time, launch count and algorithmic bytes (each input and output
 * byte counted once) accumulated for one kernel class since timing was on. */
int vr_ctx_timing(vr_ctx* ctx, int kernel_class, double* total_ms, uint64_t* launches, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* VOLRUS_H */
