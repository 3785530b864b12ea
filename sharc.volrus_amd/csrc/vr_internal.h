// Host-side objects behind the C ABI handles.
#pragma once

#include <cstdarg>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/volrus.h"
#include "vr_kernels.h"

#define NANO_HEADER_BYTES 320

// A growable device allocation reused across calls, or (borrow) a fixed
// range of memory owned elsewhere, which ensure() never reallocates: it
// fails with hipErrorOutOfMemory when the range is too small.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  hipError_t ensure(size_t n);
  void borrow(void* ptr, size_t n) {
    p = ptr;
    bytes = n;
    owned = false;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Leaf-set lookup arrays shared by every grid with the same leaf origins
// (an advected grid shares its input's).
struct Topo {
  int cap = 0;        // capacity of origins / nbr
  int count = 0;      // host copy of the live leaf count
  bool has_empty = false;  // some leaf has no active voxel
  int* d_count = nullptr;
  int4* d_origins = nullptr;
  std::vector<int4> roots;  // sorted internal-node origins (host copy)
  int4* d_roots = nullptr;
  int* d_child = nullptr;
  int* d_nbr = nullptr;
  // Second count and neighbour table: an in-place compaction writes the new
  // ones there while it reads the old, and the pointers are then swapped.
  int* d_count_alt = nullptr;
  int* d_nbr_alt = nullptr;
  Topo() = default;
  Topo(const Topo&) = delete;
  Topo& operator=(const Topo&) = delete;
  ~Topo();
};

enum ScratchSlot {
  SCR_POS = 0, SCR_SUMS, SCR_SUMS2, SCR_TOTAL, SCR_KEEP, SCR_CROSS, SCR_ANY, SCR_WORDS, SCR_COORDS,
  SCR_MC_KEEP, SCR_MC_POS, SCR_MC_TILES, SCR_MC_TLEAF, SCR_MC_TTILE, SCR_MC_META, SCR_MC_TV, SCR_MC_TT,
  SCR_MC_VB, SCR_MC_TB, SCR_NEWCOUNT, SCR_TMP_LEAVES, SCR_ACTIVE, SCR_PING, SCR_ROLL_LEAVES, SCR_ROLL_MESH,
  SCR_COUNT
};

// Device scratch of the marching-cubes passes (reused across calls).
struct MCMulti {
  DevBuf views, leaf_base, keep, pos, sums, sums2, totals, tile_base, tiles, tanchor, tleaf, ttile, cmask, meta, tv, tt,
      vb, tb, ranges;
  // Mixed-cell lists handed from the count pass to the emit pass (masks):
  // 4 bytes per mixed cell (mcell, sized by the caller), offset and count
  // per tile (mtile), the 64-bit allocation counter (mctr).
  DevBuf mcell, mtile, mctr;
};

// Kernel classes timed with HIP events when timing is on.
// TC_MC: the marching-cubes count pass; TC_MC_EMIT: its emit pass.
enum TimingClass { TC_ADVECT = 0, TC_CROSSINGS, TC_REBUILD, TC_COMPACT, TC_MC, TC_GRAPH, TC_MC_EMIT, TC_COUNT };

// bytes = per_leaf * (leaf count) + fixed.  A rollout keeps leaf counts on
// the device; its records carry the segment (and, for mesh writes, the mesh)
// they belong to and get their bytes once the counts are read back.
struct TimingRec {
  hipEvent_t a, b;
  int cls;
  double bytes;
  double per_leaf, fixed;
  int seg, mesh;
  int launches;  // launches between a and b (a run of back-to-back launches of one class)
};

struct vr_ctx {
  int device = 0;
  int cus = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  DevBuf scratch[SCR_COUNT];
  MCMulti mc;
  bool timing = false;
  std::vector<TimingRec> pending;
  std::vector<hipEvent_t> event_pool;
  int cross_gen = 0;   // generation of the last crossings pass (its "any crossing" mark)
  const void* cross_any = nullptr;  // the mark's buffer that generation counts on
  int roll_seg = -1;   // rollout segment of the launches being recorded (-1: none)
  int roll_mesh = -1;  // rollout mesh index of the launches being recorded
  // An open run of back-to-back launches of class group_cls timed by one
  // event pair (timing_group_begin/end): its launches record no events of
  // their own, only their count and bytes.
  int group_cls = -1, group_n = 0;
  hipEvent_t group_a = nullptr;
  double group_per_leaf = 0.0, group_fixed = 0.0, group_leaves = 0.0;
  double t_ms[TC_COUNT] = {};
  double t_bytes[TC_COUNT] = {};
  uint64_t t_n[TC_COUNT] = {};
  ~vr_ctx();
};

// Records an event pair around a kernel launch when ctx->timing is set.
struct TimedLaunch {
  vr_ctx* ctx;
  int cls;
  double per_leaf, fixed, leaves;
  bool mesh;  // fixed = bytes of the mesh written (known later in a rollout)
  hipEvent_t a = nullptr;
  TimedLaunch(vr_ctx* c, int k, double per_leaf_bytes, double fixed_bytes, double leaf_count, bool mesh_bytes = false);
  ~TimedLaunch();
};
// Opens / closes a run of back-to-back launches of one class timed as a
// whole (no event packets between them); no-ops when timing is off.
void timing_group_begin(vr_ctx* ctx, int cls);
void timing_group_end(vr_ctx* ctx);


struct vr_grid {
  std::shared_ptr<Topo> topo;
  NanoLeaf* d_leaves = nullptr;
  int cap = 0;
  uint32_t grid_class = 0;
  double voxel_size = 1.0;
  float background = 0.0f;
  double mat[4][4];
  double inv[4][4];
  vr_grid() = default;
  vr_grid(const vr_grid&) = delete;
  vr_grid& operator=(const vr_grid&) = delete;
  ~vr_grid();
};

struct vr_velocity {
  int kind = 0;
  double a[6] = {0, 0, 0, 0, 0, 0};
  double period = 0.0;
  const vr_grid* g[3] = {nullptr, nullptr, nullptr};
  float* d_dense = nullptr;
  int n = 0;
  double h = 1.0;
  vr_velocity() = default;
  vr_velocity(const vr_velocity&) = delete;
  vr_velocity& operator=(const vr_velocity&) = delete;
  ~vr_velocity();
};

struct vr_mesh {
  uint64_t nv = 0, nt = 0;
  uint64_t cap_v = 0, cap_t = 0;
  double* d_v = nullptr;
  uint32_t* d_t = nullptr;
  vr_mesh() = default;
  vr_mesh(const vr_mesh&) = delete;
  vr_mesh& operator=(const vr_mesh&) = delete;
  ~vr_mesh();
};

struct MCScratch {};

int vr_fail(vr_ctx* ctx, int code, const char* fmt, ...);
int vr_hip_fail(vr_ctx* ctx, hipError_t e, const char* where);
GridView make_view(const vr_grid* g);
VelView make_vel_view(const vr_velocity* v);
int build_topology(vr_ctx* ctx, vr_grid* g, std::shared_ptr<Topo> t);
// exact = false: the host leaf count stays an upper bound (no readback; the
// device count is exact) -- rollouts only, which read the counts back once.
// tail_clear: the caller already cleared the flags past the device count.
// count_copy (optional): also receives the new device count.
int compact_into(vr_ctx* ctx, vr_grid* g, const NanoLeaf* src, const int* keep, bool has_empty_after,
                 bool exact = true, bool tail_clear = false, int* count_copy = nullptr);
int advect_into(vr_ctx* ctx, const vr_grid* in, const vr_velocity* vel, double dt, double t, int scheme,
                vr_grid* o);
// *any == *gen after the pass when some voxel is a crossing.
int crossings_into(vr_ctx* ctx, const vr_grid* g, uint64_t** cross, int** any, int* gen);
// deferred: no host synchronisation; the leaf count is the host upper bound,
// the mesh is written up to its capacity and its counts are copied to
// d_counts[0..1] (vertices, triangles) instead of being read back.
int mc_run(vr_ctx* ctx, const vr_grid* g, float iso, MCScratch* s, vr_mesh* m, bool deferred = false,
           int* d_counts = nullptr);
// Resolves the device address of the allocation-free marching-cubes scratch
// on `device` (the current device) once per device; vr_ctx_create calls it.
hipError_t mc_noalloc_setup(int device);
int mc_tiles_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot,
                   const GridView* single = nullptr);
int mc_tiles_finish(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot,
                    int tcap);
// masks: the views carry sign masks for `iso` (GridView::signs), which the
// cells are classified from instead of the values; the count pass then also
// writes each tile's mixed cells to s.mcell for the emit pass, each of its
// waves into an equal share (*waves_out: the count of waves), and the
// largest share a wave needed, in cells, to s.mctr.
int mc_count_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, long long ltot, int tcap, float iso,
                   bool masks = false, long long* waves_out = nullptr);
// Timed (class TC_MC_EMIT) with algorithmic bytes per leaf (2128, the leaf,
// or with masks its 2048 value bytes and 64-byte sign mask) * leaf_count +
// mesh_bytes (24 per vertex + 12 per triangle written).
int mc_emit_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, int G, int tcap, float iso, double* verts,
                  uint32_t* tris, double leaf_count, double mesh_bytes, long long vcap = -1, long long tricap = -1,
                  int* counts_out = nullptr, bool masks = false);
// active_out (optional): receives += the active voxels after the rebuild.
int rebuild_impl(vr_ctx* ctx, vr_grid* g, float half_width, NanoLeaf* tmp, bool exact = true,
                 int* count_copy = nullptr, unsigned long long* active_out = nullptr);
// zero = false: *slot is already 0.
int count_active_into(vr_ctx* ctx, const vr_grid* g, unsigned long long* slot, bool zero = true);
void copy_header_into(vr_grid* dst, const vr_grid* src);
double vr_time_factor(double period, double t);
void vr_stage_factors(double period, double t, double dt, int scheme, double tf[3]);
