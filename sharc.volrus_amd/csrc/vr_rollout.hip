// Multi-step rollout (advect, periodic band rebuild, periodic marching
// cubes) on one grid with two ping-pong leaf buffers, and the per-kernel
// HIP-event timing the benchmark reads.
//
// The reference has no rollout driver (nothing in the crate chains
// advect -> rebuild -> MC); this composes its three tools in that order.
#include <cstring>

#include "vr_internal.h"

// ------------------------------------------------------------------ event timing
TimedLaunch::TimedLaunch(vr_ctx* c, int k, double per_leaf_bytes, double fixed_bytes, double leaf_count,
                         bool mesh_bytes)
    : ctx(c), cls(k), per_leaf(per_leaf_bytes), fixed(fixed_bytes), leaves(leaf_count), mesh(mesh_bytes) {
  if (!ctx || !ctx->timing) return;
  if (ctx->group_cls == k) {
    // inside a timed run: counted, not bracketed
    ctx->group_n += 1;
    ctx->group_per_leaf = per_leaf_bytes;
    ctx->group_fixed = fixed_bytes;
    ctx->group_leaves = leaf_count;
    return;
  }
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
  ctx->pending.push_back({a, b, cls, per_leaf * leaves + fixed, per_leaf, fixed, ctx->roll_seg,
                          mesh ? ctx->roll_mesh : -1, 1});
}

static hipEvent_t pool_event(vr_ctx* ctx) {
  hipEvent_t ev = nullptr;
  if (ctx->event_pool.empty()) {
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
  } else {
    ev = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  }
  return ev;
}

void timing_group_begin(vr_ctx* ctx, int cls) {
  if (!ctx->timing || ctx->group_cls >= 0) return;
  ctx->group_a = pool_event(ctx);
  if (!ctx->group_a) return;
  (void)hipEventRecord(ctx->group_a, ctx->stream);
  ctx->group_cls = cls;
  ctx->group_n = 0;
}

void timing_group_end(vr_ctx* ctx) {
  if (ctx->group_cls < 0) return;
  const int cls = ctx->group_cls, n = ctx->group_n;
  ctx->group_cls = -1;
  if (n == 0) {
    ctx->event_pool.push_back(ctx->group_a);
    return;
  }
  hipEvent_t b = pool_event(ctx);
  if (!b) return;
  (void)hipEventRecord(b, ctx->stream);
  // bytes per launch (per_leaf, fixed) times the launches of the run
  ctx->pending.push_back({ctx->group_a, b, cls, n * (ctx->group_per_leaf * ctx->group_leaves + ctx->group_fixed),
                          n * ctx->group_per_leaf, n * ctx->group_fixed, ctx->roll_seg, -1, n});
}

static void drain_timing(vr_ctx* ctx) {
  for (const TimingRec& r : ctx->pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      ctx->t_ms[r.cls] += ms;
      ctx->t_bytes[r.cls] += r.bytes;
      ctx->t_n[r.cls] += (uint64_t)r.launches;
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

__global__ void k_zero_u64(unsigned long long* p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0ull;
}

// *out += popcount of the active masks of leaves [0, *count).
__global__ __launch_bounds__(256) void k_active_sum(const NanoLeaf* __restrict__ leaves, const int* __restrict__ count,
                                                    unsigned long long* out) {
  const int n = *count * 8;
  unsigned int c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c += (unsigned int)__popcll(leaves[i >> 3].mask[i & 7]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  __shared__ unsigned int part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

// Device-side active-voxel count of g into *slot (no host synchronisation).
int count_active_into(vr_ctx* ctx, const vr_grid* g, unsigned long long* slot, bool zero) {
  const int cap = g->topo->cap < 1 ? 1 : g->topo->cap;
  size_t blocks = ((size_t)cap * 8 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (zero) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, slot);
  hipLaunchKernelGGL(k_active_sum, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, g->d_leaves, g->topo->d_count,
                     slot);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "active count");
}

__global__ void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Device copy of a grid's lookup arrays, so the copy can be compacted in place.
static int clone_topology(vr_ctx* ctx, const Topo& s, std::shared_ptr<Topo>* out) {
  hipError_t e;
  std::shared_ptr<Topo> t(new Topo());
  t->cap = s.cap;
  t->count = s.count;
  t->has_empty = s.has_empty;
  t->roots = s.roots;
  const size_t cap = (size_t)(s.cap < 1 ? 1 : s.cap), R = s.roots.size() ? s.roots.size() : 1;
  if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess ||
      (e = hipMalloc(&t->d_origins, sizeof(int4) * cap)) != hipSuccess ||
      (e = hipMalloc(&t->d_roots, sizeof(int4) * R)) != hipSuccess ||
      (e = hipMalloc(&t->d_child, sizeof(int) * 4096 * R)) != hipSuccess ||
      (e = hipMalloc(&t->d_nbr, sizeof(int) * VR_NBR * cap)) != hipSuccess ||
      (e = hipMalloc(&t->d_count_alt, sizeof(int))) != hipSuccess ||
      (e = hipMalloc(&t->d_nbr_alt, sizeof(int) * VR_NBR * cap)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMemcpyAsync(t->d_count, s.d_count, sizeof(int), hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(t->d_origins, s.d_origins, sizeof(int4) * cap, hipMemcpyDeviceToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(t->d_roots, s.d_roots, sizeof(int4) * R, hipMemcpyDeviceToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(t->d_child, s.d_child, sizeof(int) * 4096 * R, hipMemcpyDeviceToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(t->d_nbr, s.d_nbr, sizeof(int) * VR_NBR * cap, hipMemcpyDeviceToDevice, ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  *out = t;
  return VR_OK;
}

extern "C" int vr_grid_clone(vr_ctx* ctx, const vr_grid* g, vr_grid** out) {
  if (!ctx || !g || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_grid> o(new vr_grid());
  copy_header_into(o.get(), g);
  int rc = clone_topology(ctx, *g->topo, &o->topo);
  if (rc != VR_OK) return rc;
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
// Grows mesh m to the largest vertex and triangle counts of meshes 0..last
// (device counts d_mc, two per mesh; read back here, host-synchronised) and,
// when mesh `last` was cut at the old capacity, meshes g again into the new
// buffers.  Also makes g's host leaf count exact.
static int rollout_fix_last_mesh(vr_ctx* ctx, vr_grid* g, float iso, vr_mesh* m, const int* d_mc, int last) {
  hipError_t e;
  std::vector<int> mc((size_t)2 * (last + 1), 0);
  int cnt = 0;
  if ((e = hipMemcpyAsync(mc.data(), d_mc, sizeof(int) * mc.size(), hipMemcpyDeviceToHost, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(&cnt, g->topo->d_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "mesh counts");
  g->topo->count = cnt;
  uint64_t mv = 0, mt = 0;
  for (int k = 0; k <= last; ++k) {
    mv = std::max<uint64_t>(mv, (uint64_t)mc[2 * k]);
    mt = std::max<uint64_t>(mt, (uint64_t)mc[2 * k + 1]);
  }
  const bool cut = (uint64_t)mc[2 * last] > m->cap_v || (uint64_t)mc[2 * last + 1] > m->cap_t;
  if (mv > m->cap_v) {
    if (m->d_v) (void)hipFree(m->d_v);
    m->d_v = nullptr;
    m->cap_v = 0;
    if ((e = hipMalloc(&m->d_v, sizeof(double) * 3 * (size_t)(mv + 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc mesh");
    m->cap_v = mv + 1;
  }
  if (mt > m->cap_t) {
    if (m->d_t) (void)hipFree(m->d_t);
    m->d_t = nullptr;
    m->cap_t = 0;
    if ((e = hipMalloc(&m->d_t, sizeof(uint32_t) * 3 * (size_t)(mt + 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc mesh");
    m->cap_t = mt + 1;
  }
  if (!cut) return VR_OK;
  MCScratch sc;
  const int seg = ctx->roll_seg, msh = ctx->roll_mesh;
  ctx->roll_seg = ctx->roll_mesh = -1;  // not part of the rollout's timed schedule
  const bool timing = ctx->timing;
  ctx->timing = false;
  const int rc = mc_run(ctx, g, iso, &sc, m);
  ctx->timing = timing;
  ctx->roll_seg = seg;
  ctx->roll_mesh = msh;
  return rc;
}

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
  if (g->topo.use_count() > 1) {
    // Own the lookup arrays so every compaction below runs in place and
    // never reallocates the leaf buffers.
    std::shared_ptr<Topo> t;
    int rc0 = clone_topology(ctx, *g->topo, &t);
    if (rc0 != VR_OK) return rc0;
    g->topo = t;
  }
  // Counts stay on the device until the end: per topology segment (between
  // rebuilds) the active voxels and the leaf count, per mesh its vertex and
  // triangle counts.  The host leaf count is an upper bound meanwhile (it
  // only sizes buffers and launch grids; kernels read the device count).
  const int nseg = 1 + (rebuild_every > 0 ? steps / rebuild_every : 0);
  const int nmesh = mesh_every > 0 ? steps / mesh_every : 0;
  std::vector<int> seg_steps((size_t)nseg, 0);
  DevBuf& slots = ctx->scratch[SCR_ACTIVE];
  DevBuf& seg_leaves = ctx->scratch[SCR_ROLL_LEAVES];
  DevBuf& mesh_counts = ctx->scratch[SCR_ROLL_MESH];
  if ((e = slots.ensure(sizeof(unsigned long long) * (size_t)nseg)) != hipSuccess ||
      (e = seg_leaves.ensure(sizeof(int) * (size_t)nseg)) != hipSuccess ||
      (e = mesh_counts.ensure(sizeof(int) * 2 * (size_t)(nmesh + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  const size_t rec0 = ctx->pending.size();  // timing records of this rollout start here
  int seg = 0, imesh = -1;
  ctx->roll_seg = 0;
  // all segment slots zeroed at once; each rebuild's compaction copies the
  // new leaf count into the next segment's slot
  int rc = VR_OK;
  if ((e = vr_fill_int(slots.as<int>(), 0, 2 * (size_t)nseg, ctx->stream)) != hipSuccess ||
      (e = vr_copy_int(g->topo->d_count, seg_leaves.as<int>(), ctx->stream)) != hipSuccess)
    rc = vr_hip_fail(ctx, e, "counts");
  if (rc == VR_OK) rc = count_active_into(ctx, g, slots.as<unsigned long long>() + seg, false);
  if (rc != VR_OK) {
    ctx->roll_seg = -1;
    return rc;
  }
  // Second leaf buffer (context scratch; ownership is swapped back at the end);
  // the advected grid shares g's topology.
  const int cap = g->topo->cap < 1 ? 1 : g->topo->cap;
  DevBuf& ping = ctx->scratch[SCR_PING];
  if ((e = ping.ensure(sizeof(NanoLeaf) * (size_t)cap)) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
  NanoLeaf* const pbuf = ping.as<NanoLeaf>();
  const size_t own_bytes = sizeof(NanoLeaf) * (size_t)(g->cap < 1 ? 1 : g->cap);
  const size_t pbuf_bytes = ping.bytes;
  NanoLeaf* other = pbuf;
  std::unique_ptr<vr_mesh> own_mesh;
  bool last_mesh_fixed = false;  // the last mesh was checked (and regrown) where it was taken
  if (!mesh && mesh_every > 0) {
    own_mesh.reset(new vr_mesh());
    mesh = own_mesh.get();
  }
  for (int s = 0; s < steps && rc == VR_OK; ++s) {
    const double t = t0 + (double)s * dt;
    // the advect launches between two rebuild or mesh passes are timed as
    // one run (one segment, so one leaf count)
    timing_group_begin(ctx, TC_ADVECT);
    {
      // The view shares g's topology only for the launch, so the
      // compactions below see it unshared and run in place.
      vr_grid out_view;
      copy_header_into(&out_view, g);
      out_view.topo = g->topo;
      out_view.d_leaves = other;
      rc = advect_into(ctx, g, vel, dt, t, scheme, &out_view);
      out_view.d_leaves = nullptr;  // not owned by the view
    }
    if (rc != VR_OK) break;
    std::swap(g->d_leaves, other);
    seg_steps[seg] += 1;
    st.advect_steps += 1;
    const bool rebuild_now = rebuild_every > 0 && (s + 1) % rebuild_every == 0;
    const bool mesh_now = mesh_every > 0 && (s + 1) % mesh_every == 0 && imesh + 1 < nmesh;
    if (g->topo->has_empty || rebuild_now || mesh_now || s + 1 == steps) timing_group_end(ctx);
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
      rc = compact_into(ctx, g, other, keep.as<int>(), false, false);
      if (rc != VR_OK) break;
    }
    if (rebuild_now) {
      const bool next = seg + 1 < nseg;
      // the rebuild writes the next segment's leaf and active counts
      rc = rebuild_impl(ctx, g, half_width, other, false, next ? seg_leaves.as<int>() + seg + 1 : nullptr,
                        next ? slots.as<unsigned long long>() + seg + 1 : nullptr);
      if (rc != VR_OK) break;
      st.rebuilds += 1;
      if (next) {
        ++seg;
        ctx->roll_seg = seg;
      }
    }
    if (mesh_now) {
      ++imesh;
      if (g->topo->count > 0) {  // an upper bound: 0 means no leaves
        ctx->roll_mesh = imesh;
        MCScratch sc;
        rc = mc_run(ctx, g, iso, &sc, mesh, true, mesh_counts.as<int>() + 2 * imesh);
        ctx->roll_mesh = -1;
      } else if ((e = vr_fill_int(mesh_counts.as<int>() + 2 * imesh, 0, 2, ctx->stream)) != hipSuccess) {
        rc = vr_hip_fail(ctx, e, "mesh counts");
      }
      if (rc != VR_OK) break;
      st.meshes += 1;
      if (imesh + 1 == nmesh && s + 1 < steps && mesh) {
        // The last mesh is taken before the last step, so the final state is
        // not the one it shows: check it now (one synchronisation) and, if it
        // was cut at the capacity, grow the buffers and mesh this state again.
        rc = rollout_fix_last_mesh(ctx, g, iso, mesh, mesh_counts.as<int>(), imesh);
        if (rc != VR_OK) break;
        last_mesh_fixed = true;
      }
    }
  }
  timing_group_end(ctx);  // a run left open by an error
  ctx->roll_seg = -1;
  ctx->roll_mesh = -1;
  // The grid keeps whichever buffer holds its data; the scratch slot takes
  // the spare one (`other` is never freed by the steps above).
  ping.p = other;
  ping.bytes = other == pbuf ? pbuf_bytes : own_bytes;
  if (g->d_leaves == pbuf) g->cap = (int)(pbuf_bytes / sizeof(NanoLeaf));
  std::vector<unsigned long long> act((size_t)nseg, 0);
  std::vector<int> leaves((size_t)nseg, 0), mc((size_t)2 * (nmesh + 1), 0);
  if (rc == VR_OK &&
      ((e = hipMemcpyAsync(act.data(), slots.p, sizeof(unsigned long long) * nseg, hipMemcpyDeviceToHost,
                           ctx->stream)) != hipSuccess ||
       (e = hipMemcpyAsync(leaves.data(), seg_leaves.p, sizeof(int) * nseg, hipMemcpyDeviceToHost, ctx->stream)) !=
           hipSuccess ||
       (e = hipMemcpyAsync(mc.data(), mesh_counts.p, sizeof(int) * mc.size(), hipMemcpyDeviceToHost, ctx->stream)) !=
           hipSuccess ||
       (e = hipStreamSynchronize(ctx->stream)) != hipSuccess))
    rc = vr_hip_fail(ctx, e, "readback");
  if (rc != VR_OK) {
    // leave the host count exact for later calls, if the device still answers
    int cnt = 0;
    if (hipMemcpy(&cnt, g->topo->d_count, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess) g->topo->count = cnt;
    return rc;
  }
  g->topo->count = leaves[seg];  // exact again
  // Algorithmic bytes of this rollout's timed launches, from the real counts.
  for (size_t i = rec0; i < ctx->pending.size(); ++i) {
    TimingRec& r = ctx->pending[i];
    if (r.seg < 0) continue;
    const double mesh_bytes = r.mesh >= 0 ? 24.0 * mc[2 * r.mesh] + 12.0 * mc[2 * r.mesh + 1] : r.fixed;
    // (for a run of launches, per_leaf and fixed already hold the run's totals)
    r.bytes = r.per_leaf * leaves[r.seg] + mesh_bytes;
  }
  for (int k = 0; k <= seg; ++k) st.voxel_updates += act[k] * (uint64_t)seg_steps[k];
  if (mesh && imesh >= 0 && last_mesh_fixed) {
    mesh->nv = (uint64_t)mc[2 * imesh];
    mesh->nt = (uint64_t)mc[2 * imesh + 1];
  } else if (mesh && imesh >= 0) {
    // Meshes past the capacity were written up to it: grow the capacity to
    // the largest count, and mesh the final state (the last mesh's state:
    // the last mesh is taken at the last step) again if it was cut.
    uint64_t mv = 0, mt = 0;
    for (int k = 0; k <= imesh; ++k) {
      mv = std::max<uint64_t>(mv, (uint64_t)mc[2 * k]);
      mt = std::max<uint64_t>(mt, (uint64_t)mc[2 * k + 1]);
    }
    mesh->nv = (uint64_t)mc[2 * imesh];
    mesh->nt = (uint64_t)mc[2 * imesh + 1];
    if (mv > mesh->cap_v || mt > mesh->cap_t) {
      if (mv > mesh->cap_v) {
        if (mesh->d_v) (void)hipFree(mesh->d_v);
        mesh->d_v = nullptr;
        mesh->cap_v = 0;
        if ((e = hipMalloc(&mesh->d_v, sizeof(double) * 3 * (size_t)(mv + 1))) != hipSuccess)
          return vr_hip_fail(ctx, e, "hipMalloc mesh");
        mesh->cap_v = mv + 1;
      }
      if (mt > mesh->cap_t) {
        if (mesh->d_t) (void)hipFree(mesh->d_t);
        mesh->d_t = nullptr;
        mesh->cap_t = 0;
        if ((e = hipMalloc(&mesh->d_t, sizeof(uint32_t) * 3 * (size_t)(mt + 1))) != hipSuccess)
          return vr_hip_fail(ctx, e, "hipMalloc mesh");
        mesh->cap_t = mt + 1;
      }
      // new buffers: mesh the final state into them (host-synchronised)
      MCScratch sc;
      rc = mc_run(ctx, g, iso, &sc, mesh);
      if (rc != VR_OK) return rc;
    }
  }
  if (mesh) {
    st.last_vertex_count = mesh->nv;
    st.last_triangle_count = mesh->nt;
  }
  st.final_leaf_count = (uint64_t)g->topo->count;
  st.final_active_count = act[seg];
  if (stats) *stats = st;
  return VR_OK;
}
