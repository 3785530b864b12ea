// Multi-step rollout (advect, periodic band rebuild, periodic marching
// cubes) on one grid with two ping-pong leaf buffers, and the per-kernel
// HIP-event timing the benchmark reads.
//
// The reference has no rollout driver (nothing in the crate chains
// advect -> rebuild -> MC); this composes its three tools in that order.
#include <cstring>

#include "vr_internal.h"

// ------------------------------------------------------------------ event timing
TimedLaunch::TimedLaunch(vr_ctx* c, int k, double b) : ctx(c), cls(k), bytes(b) {
  if (!ctx || !ctx->timing) return;
  if (ctx->event_pool.empty()) {
    if (hipEventCreate(&a) != hipSuccess) {
      a = nullptr;
      return;
    }
  } else {
    a = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  }
  (void)hipEventRecord(a, ctx->stream);
}

TimedLaunch::~TimedLaunch() {
  if (!a) return;
  hipEvent_t b = nullptr;
  if (ctx->event_pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  }
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.push_back({a, b, cls, bytes});
}

static void drain_timing(vr_ctx* ctx) {
  for (const TimingRec& r : ctx->pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      ctx->t_ms[r.cls] += ms;
      ctx->t_bytes[r.cls] += r.bytes;
      ctx->t_n[r.cls] += 1;
    }
    ctx->event_pool.push_back(r.a);
    ctx->event_pool.push_back(r.b);
  }
  ctx->pending.clear();
}

extern "C" int vr_ctx_set_timing(vr_ctx* ctx, int enable) {
  if (!ctx) return VR_ERR_INVALID_ARGUMENT;
  drain_timing(ctx);
  ctx->timing = enable != 0;
  for (int i = 0; i < TC_COUNT; ++i) {
    ctx->t_ms[i] = 0.0;
    ctx->t_bytes[i] = 0.0;
    ctx->t_n[i] = 0;
  }
  return VR_OK;
}

extern "C" int vr_ctx_timing(vr_ctx* ctx, int kernel_class, double* total_ms, uint64_t* launches, double* bytes) {
  if (!ctx || kernel_class < 0 || kernel_class >= TC_COUNT) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad class");
  drain_timing(ctx);
  if (total_ms) *total_ms = ctx->t_ms[kernel_class];
  if (launches) *launches = ctx->t_n[kernel_class];
  if (bytes) *bytes = ctx->t_bytes[kernel_class];
  return VR_OK;
}

// ------------------------------------------------------------------ active count
int active_count(vr_ctx* ctx, const vr_grid* g, uint64_t* out) {
  uint64_t lc;
  return vr_grid_counts(ctx, g, &lc, out);
}

__global__ void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int vr_grid_clone(vr_ctx* ctx, const vr_grid* g, vr_grid** out) {
  if (!ctx || !g || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_grid> o(new vr_grid());
  copy_header_into(o.get(), g);
  o->topo = g->topo;
  o->cap = g->topo->count;
  const size_t n = (size_t)(o->cap ? o->cap : 1);
  if ((e = hipMalloc(&o->d_leaves, sizeof(NanoLeaf) * n)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  const size_t nu4 = (size_t)g->topo->count * (VR_LEAF_BYTES / 16);
  if (nu4) {
    size_t blocks = (nu4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy_u4, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const uint4*>(g->d_leaves), reinterpret_cast<uint4*>(o->d_leaves), nu4);
    if ((e = hipGetLastError()) != hipSuccess) return vr_hip_fail(ctx, e, "copy");
  }
  *out = o.release();
  return VR_OK;
}

extern "C" int vr_mesh_create(vr_ctx* ctx, vr_mesh** out) {
  if (!ctx || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = new vr_mesh();
  return VR_OK;
}

// ------------------------------------------------------------------ rollout
extern "C" int vr_rollout(vr_ctx* ctx, vr_grid* g, const vr_velocity* vel, double t0, double dt, int scheme,
                          int steps, int rebuild_every, float half_width, int mesh_every, float iso, vr_mesh* mesh,
                          vr_rollout_stats* stats) {
  if (!ctx || !g || !vel || steps < 0) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad rollout arguments");
  if (scheme != VR_SCHEME_RK1 && scheme != VR_SCHEME_RK3)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "scheme must be 1 or 3");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  vr_rollout_stats st;
  std::memset(&st, 0, sizeof(st));
  uint64_t active = 0;
  int rc = active_count(ctx, g, &active);
  if (rc != VR_OK) return rc;
  // Second leaf buffer; the advected grid shares g's topology.
  NanoLeaf* other = nullptr;
  const int cap = g->topo->cap < 1 ? 1 : g->topo->cap;
  if ((e = hipMalloc(&other, sizeof(NanoLeaf) * (size_t)cap)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  std::unique_ptr<vr_mesh> own_mesh;
  if (!mesh && mesh_every > 0) {
    own_mesh.reset(new vr_mesh());
    mesh = own_mesh.get();
  }
  for (int s = 0; s < steps && rc == VR_OK; ++s) {
    const double t = t0 + (double)s * dt;
    vr_grid out_view;
    copy_header_into(&out_view, g);
    out_view.topo = g->topo;
    out_view.d_leaves = other;
    rc = advect_into(ctx, g, vel, dt, t, scheme, &out_view);
    out_view.d_leaves = nullptr;  // not owned by the view
    if (rc != VR_OK) break;
    std::swap(g->d_leaves, other);
    st.voxel_updates += active;
    st.advect_steps += 1;
    if (g->topo->has_empty) {
      // Inputs with leaves that hold no active voxel: the first step drops them.
      DevBuf& keep = ctx->scratch[SCR_KEEP];
      if ((e = keep.ensure(sizeof(int) * (size_t)(g->topo->count + 1))) != hipSuccess) {
        rc = vr_hip_fail(ctx, e, "scratch");
        break;
      }
      if ((e = vr_launch_nonempty(g->d_leaves, g->topo->d_count, g->topo->cap, keep.as<int>(), ctx->stream)) !=
          hipSuccess) {
        rc = vr_hip_fail(ctx, e, "nonempty");
        break;
      }
      std::swap(g->d_leaves, other);
      rc = compact_into(ctx, g, other, keep.as<int>(), false);
      if (rc != VR_OK) break;
    }
    if (rebuild_every > 0 && (s + 1) % rebuild_every == 0) {
      rc = rebuild_impl(ctx, g, half_width, other);
      if (rc != VR_OK) break;
      rc = active_count(ctx, g, &active);
      st.rebuilds += 1;
      if (rc != VR_OK) break;
    }
    if (mesh_every > 0 && (s + 1) % mesh_every == 0) {
      if (g->topo->count > 0) {
        MCScratch sc;
        rc = mc_run(ctx, g, iso, &sc, mesh);
        if (rc != VR_OK) break;
      } else {
        mesh->nv = mesh->nt = 0;
      }
      st.meshes += 1;
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(other);
  if (rc != VR_OK) return rc;
  if (mesh) {
    st.last_vertex_count = mesh->nv;
    st.last_triangle_count = mesh->nt;
  }
  st.final_leaf_count = (uint64_t)g->topo->count;
  st.final_active_count = active;
  if (stats) *stats = st;
  return VR_OK;
}
