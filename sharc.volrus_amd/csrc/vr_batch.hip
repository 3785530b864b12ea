// Batched advection + marching cubes over many level sets.
//
// No reference counterpart (the reference handles one grid per call).  The
// leaves of all grids form one global worklist, cut into chunks of up to
// `leaves_per_grab` consecutive leaves of one grid.  One persistent launch
// of `waves_per_cu` waves per CU advects them: the chunk list is cut into 8
// contiguous ranges, one per XCD (blocks are dispatched to the 8 XCDs
// round-robin by block index, so block b serves range b % 8), and blocks
// take chunks from their range's atomic counter, so neighbouring leaves,
// whose halos overlap, stay in one XCD's L2.  The leaf code is specialised
// for the velocity kind the grids share, with one wave per leaf once the
// batch has enough leaves to fill the device.  Marching cubes then runs over the
// concatenated tiles of all grids (vr_mc.hip).  The whole step can be
// captured once into a hipGraph and replayed; `vr_batch_autotune` picks
// leaves-per-grab and waves-per-CU by timing the advect launch.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "vr_advect.h"
#include "vr_internal.h"

#define VK_MIXED (-1)  // grids of one batch with different velocity kinds

struct vr_batch {
  int G = 0;
  // The batch's own copies of the input leaves (sharing the inputs' lookup
  // arrays, which no later call on an input changes in place while they are
  // shared), so the inputs may be modified or destroyed after creation.
  std::vector<std::unique_ptr<vr_grid>> in;
  std::vector<const vr_velocity*> vel;
  std::vector<std::unique_ptr<vr_grid>> out;  // advected grids (own leaves, share topology)
  std::vector<int> leaf_base;                 // G + 1
  long long ltot = 0;
  int lpg = 4, wpc = 8;                       // leaves per grab, waves per CU
  int vk = VK_MIXED;                          // velocity kind shared by all grids
  int cus = 256;
  DevBuf d_in, d_out_views, d_vel, d_outs, d_leaf_base, d_counters, d_ranges, d_dts, d_chunks;
  DevBuf d_signs;  // 16 words per leaf (global leaf index): sign masks of the outputs
  int chunk_lpg = 0;  // leaves per chunk the chunk table was built for
  int max_range = 0;  // most chunks in one XCD range
  bool has_dts = false;
  std::vector<double> dts;
  MCMulti mc;
  int tcap = 0;
  bool tiles_ready = false;  // mc holds the tile lists (fixed: advection keeps every leaf set)
  long long vcap = 0, tricap = 0;
  double* d_verts = nullptr;
  uint32_t* d_tris = nullptr;
  std::vector<int> ranges_host;  // 4 per grid after a step
  // Mesh capacities come from a host-synchronised sizing run; they hold for
  // the step parameters of that run (s_*) and the per-grid dts of that time.
  bool sized = false;
  double s_dt = 0, s_t = 0;
  int s_scheme = 0;
  float s_iso = 0;
  // The captured step and the parameters it was captured with.  Any change
  // to a buffer the graph reads or writes (chunk table, mesh buffers, dts)
  // drops it (batch_drop_graph), so a replay never sees a freed buffer.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double g_dt = 0, g_t = 0;
  int g_scheme = 0, g_lpg = 0, g_wpc = 0;
  float g_iso = 0;
  std::vector<std::unique_ptr<vr_mesh>> loop_mesh;  // per-grid meshes of the loop mode
  ~vr_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_verts) (void)hipFree(d_verts);
    if (d_tris) (void)hipFree(d_tris);
  }
};

static void batch_drop_graph(vr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// ------------------------------------------------------------------ persistent advect
// VK = the velocity kind shared by every grid of the batch (the leaf code is
// specialised for it), or VK_MIXED when kinds differ (a per-leaf switch);
// NW waves per leaf over its compacted active list.

// Leaf li of a grid with velocity v from its loads ld; VK_MIXED switches on
// the grid's velocity kind.
template <int VK, int SCHEME, int NW, class LDS>
__device__ __forceinline__ void batch_leaf(const GridView& in, const VelView& v, NanoLeaf* out, int li,
                                           const AdvectLoads<NW>& ld, LDS& s, int tid, double dt, const double tf[3],
                                           uint32_t* sg, float iso) {
  if constexpr (VK != VK_MIXED) {
    advect_leaf_loaded<VK, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso);
  } else {
    switch (v.kind) {
      case VK_UNIFORM: advect_leaf_loaded<VK_UNIFORM, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso); break;
      case VK_RIGID: advect_leaf_loaded<VK_RIGID, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso); break;
      case VK_VORTEX: advect_leaf_loaded<VK_VORTEX, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso); break;
      case VK_DENSE: advect_leaf_loaded<VK_DENSE, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso); break;
      default: advect_leaf_loaded<VK_GRIDS3, SCHEME, NW, true>(in, v, out, li, ld, s, tid, dt, tf, sg, iso); break;
    }
  }
}

template <int VK, int SCHEME, int NW, bool ONE, class LDS>
__device__ __forceinline__ void batch_chunk(const GridView* ins, const VelView* vels, NanoLeaf* const* outs,
                                            const int4& ck, LDS& s, int tid, const double* dts, double dt,
                                            const double tf[3], uint32_t* signs, const int* leaf_base, float iso) {
  const GridView in = ins[ck.x];
  const VelView v = vels[ck.x];
  // Typed global where loaded (kernel parameter types stay unqualified: the
  // host pass has no address spaces and must name the same kernel).
  NanoLeaf* const out = (VR_GLOBAL NanoLeaf*)reinterpret_cast<const uintptr_t*>(outs)[ck.x];
  const double gdt = dts ? dts[ck.x] : dt;
  uint32_t* const sg = signs + (size_t)leaf_base[ck.x] * 16;
  // ONE: chunks of one leaf, so no loop (nothing is kept live across leaves).
  const int li_end = ONE ? ck.y + 1 : ck.y + ck.z;
  for (int li = ck.y; li < li_end; ++li)
    batch_leaf<VK, SCHEME, NW>(in, v, out, li, advect_issue<NW>(in.leaves, in.nbr, li, tid), s, tid, gdt, tf,
                               sg + li * 16, iso);
}

// persistent != 0: `gridDim.x` blocks loop, taking chunks from their XCD's
// counter.  persistent == 0: one chunk per block, block b taking chunk
// b / 8 of XCD b % 8's range (the grid covers 8 x the longest range).  ONE:
// one chunk per block and chunks of one leaf, so chunk c is global leaf c:
// its grid is found by a search of leaf_base (G + 1 entries, in the scalar
// cache) instead of a load from the chunk table, which would start each
// block's chain of dependent loads with a miss.
template <int VK, int SCHEME, int NW, bool ONE>
__global__ __launch_bounds__(64 * NW) void k_batch_advect(const GridView* ins, const VelView* vels,
                                                          NanoLeaf* const* outs, const int4* chunks, int* counters,
                                                          const int* ranges, int persistent, const double* dts,
                                                          double dt, double tf0, double tf1, double tf2,
                                                          uint32_t* signs, const int* leaf_base, int G, float iso) {
  __shared__ AdvectLDS<VK> s;
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int xcd = blockIdx.x & 7;
  const int lo = ranges[xcd], hi = ranges[xcd + 1];
  const double tf[3] = {tf0, tf1, tf2};
  if (ONE) {
    const int c = lo + (int)(blockIdx.x >> 3);
    if (c < hi) {
      int g0 = 0, g1 = G;  // leaf_base[g0] <= c < leaf_base[g1]
      while (g1 - g0 > 1) {
        const int mid = (g0 + g1) >> 1;
        if (leaf_base[mid] <= c) g0 = mid; else g1 = mid;
      }
      batch_chunk<VK, SCHEME, NW, true>(ins, vels, outs, make_int4(g0, c - leaf_base[g0], 1, 0), s, tid, dts, dt, tf,
                                        signs, leaf_base, iso);
    }
    return;
  }
  if (!persistent) {
    const int c = lo + (int)(blockIdx.x >> 3);
    if (c < hi) batch_chunk<VK, SCHEME, NW, false>(ins, vels, outs, chunks[c], s, tid, dts, dt, tf, signs, leaf_base, iso);
    return;
  }
  // Chunk indices read from LDS are block-uniform; readfirstlane tells the
  // compiler so, and the chunk and grid descriptors load as scalars.
  if (tid == 0) s_item = atomicAdd(&counters[xcd], 1);
  __syncthreads();
  int c = __builtin_amdgcn_readfirstlane(lo + s_item);
  __syncthreads();
  while (c < hi) {
    // Grab the next chunk now; its latency overlaps this chunk's loads.
    int next = 0;
    if (tid == 0) next = atomicAdd(&counters[xcd], 1);
    batch_chunk<VK, SCHEME, NW, false>(ins, vels, outs, chunks[c], s, tid, dts, dt, tf, signs, leaf_base, iso);
    if (tid == 0) s_item = next;
    __syncthreads();
    c = __builtin_amdgcn_readfirstlane(lo + s_item);
    __syncthreads();
  }
}

template <int VK, int SCHEME, int NW, bool ONE>
static void launch_batch_k(int blocks, hipStream_t st, const GridView* ins, const VelView* vels,
                           NanoLeaf* const* outs, const int4* chunks, int* counters, const int* ranges, int persistent,
                           const double* dts, double dt, const double tf[3], uint32_t* signs,
                           const int* leaf_base, int G, float iso) {
  hipLaunchKernelGGL((k_batch_advect<VK, SCHEME, NW, ONE>), dim3(blocks), dim3(64 * NW), 0, st, ins, vels, outs,
                     chunks, counters, ranges, persistent, dts, dt, tf[0], tf[1], tf[2], signs, leaf_base, G, iso);
}

template <int VK, int SCHEME>
static void launch_batch_t(int nw, bool one, int blocks, hipStream_t st, const GridView* ins, const VelView* vels,
                           NanoLeaf* const* outs, const int4* chunks, int* counters, const int* ranges, int persistent,
                           const double* dts, double dt, const double tf[3], uint32_t* signs,
                           const int* leaf_base, int G, float iso) {
  if (nw == 1) {
    if (one)
      launch_batch_k<VK, SCHEME, 1, true>(blocks, st, ins, vels, outs, chunks, counters, ranges, persistent, dts, dt, tf, signs, leaf_base, G, iso);
    else
      launch_batch_k<VK, SCHEME, 1, false>(blocks, st, ins, vels, outs, chunks, counters, ranges, persistent, dts, dt,
                                           tf, signs, leaf_base, G, iso);
  } else {
    if (one)
      launch_batch_k<VK, SCHEME, 4, true>(blocks, st, ins, vels, outs, chunks, counters, ranges, persistent, dts, dt, tf, signs, leaf_base, G, iso);
    else
      launch_batch_k<VK, SCHEME, 4, false>(blocks, st, ins, vels, outs, chunks, counters, ranges, persistent, dts, dt,
                                           tf, signs, leaf_base, G, iso);
  }
}

template <int SCHEME>
static void launch_batch_s(int vk, int nw, bool one, int blocks, hipStream_t st, const GridView* ins, const VelView* vels,
                           NanoLeaf* const* outs, const int4* chunks, int* counters, const int* ranges,
                           int persistent, const double* dts, double dt, const double tf[3], uint32_t* signs,
                           const int* leaf_base, int G, float iso) {
#define VR_BK(K) \
  launch_batch_t<K, SCHEME>(nw, one, blocks, st, ins, vels, outs, chunks, counters, ranges, persistent, dts, dt, tf, signs, leaf_base, G, iso)
  switch (vk) {
    case VK_UNIFORM: VR_BK(VK_UNIFORM); break;
    case VK_RIGID: VR_BK(VK_RIGID); break;
    case VK_VORTEX: VR_BK(VK_VORTEX); break;
    case VK_DENSE: VR_BK(VK_DENSE); break;
    case VK_GRIDS3: VR_BK(VK_GRIDS3); break;
    default: VR_BK(VK_MIXED); break;
  }
#undef VR_BK
}

__global__ void k_batch_reset(int* counters) {
  if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

// ------------------------------------------------------------------ API
// The global worklist as chunks of up to `lpg` consecutive leaves of one
// grid, (grid, first leaf, leaf count, 0), in grid order; cut into 8
// contiguous ranges of chunks, one per XCD.  Rebuilt when lpg changes
// (never inside a graph capture: the step builds it before capturing).
static int batch_chunks(vr_ctx* ctx, vr_batch* b) {
  if (b->chunk_lpg == b->lpg && b->d_chunks.p) return VR_OK;
  batch_drop_graph(b);  // the table (and maybe its buffer) changes
  std::vector<int4> ch;
  for (int g = 0; g < b->G; ++g) {
    const int n = b->leaf_base[g + 1] - b->leaf_base[g];
    for (int f = 0; f < n; f += b->lpg) ch.push_back(make_int4(g, f, n - f < b->lpg ? n - f : b->lpg, 0));
  }
  const long long nc = (long long)ch.size();
  int ranges[9];
  for (int x = 0; x <= 8; ++x) ranges[x] = (int)((nc * x) / 8);
  b->max_range = 0;
  for (int x = 0; x < 8; ++x) b->max_range = std::max(b->max_range, ranges[x + 1] - ranges[x]);
  hipError_t e;
  if ((e = b->d_chunks.ensure(sizeof(int4) * (ch.size() ? ch.size() : 1))) != hipSuccess ||
      (e = b->d_ranges.ensure(sizeof(int) * 9)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  if ((ch.size() && (e = hipMemcpyAsync(b->d_chunks.p, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice,
                                        ctx->stream)) != hipSuccess) ||
      (e = hipMemcpyAsync(b->d_ranges.p, ranges, sizeof(ranges), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpy");
  b->chunk_lpg = b->lpg;
  return VR_OK;
}

extern "C" int vr_batch_create(vr_ctx* ctx, vr_grid* const* grids, const vr_velocity* const* vels, int count,
                               vr_batch** out) {
  if (!ctx || !grids || !vels || count <= 0 || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad batch");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_batch> b(new vr_batch());
  b->G = count;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
    b->cus = prop.multiProcessorCount;
  b->leaf_base.assign((size_t)count + 1, 0);
  std::vector<GridView> vin, vout;
  std::vector<VelView> vv;
  std::vector<NanoLeaf*> outs;
  for (int i = 0; i < count; ++i) {
    vr_grid* g = grids[i];
    const vr_velocity* v = vels[i];
    if (!g || !v) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null grid or velocity %d", i);
    if (g->topo->has_empty) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "grid %d has leaves without active voxels", i);
    const int n = g->topo->count;
    std::unique_ptr<vr_grid> c(new vr_grid());
    copy_header_into(c.get(), g);
    c->topo = g->topo;
    c->cap = n;
    if ((e = hipMalloc(&c->d_leaves, sizeof(NanoLeaf) * (size_t)(n ? n : 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc");
    if (n && (e = hipMemcpyAsync(c->d_leaves, g->d_leaves, sizeof(NanoLeaf) * (size_t)n, hipMemcpyDeviceToDevice,
                                 ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMemcpyAsync");
    b->vel.push_back(v);
    std::unique_ptr<vr_grid> o(new vr_grid());
    copy_header_into(o.get(), g);
    o->topo = g->topo;
    o->cap = n;
    if ((e = hipMalloc(&o->d_leaves, sizeof(NanoLeaf) * (size_t)(o->cap ? o->cap : 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc");
    vin.push_back(make_view(c.get()));
    b->in.push_back(std::move(c));
    vout.push_back(make_view(o.get()));
    vv.push_back(make_vel_view(v));
    outs.push_back(o->d_leaves);
    b->out.push_back(std::move(o));
    b->leaf_base[i + 1] = b->leaf_base[i] + g->topo->count;
  }
  b->ltot = b->leaf_base[count];
  b->vk = vels[0]->kind;
  for (int i = 1; i < count; ++i)
    if (vels[i]->kind != b->vk) b->vk = VK_MIXED;
  if (b->ltot > 0x7fffffff / 7) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "batch too large");
  if ((e = b->d_in.ensure(sizeof(GridView) * count)) != hipSuccess ||
      (e = b->d_out_views.ensure(sizeof(GridView) * count)) != hipSuccess ||
      (e = b->d_vel.ensure(sizeof(VelView) * count)) != hipSuccess ||
      (e = b->d_outs.ensure(sizeof(NanoLeaf*) * count)) != hipSuccess ||
      (e = b->d_leaf_base.ensure(sizeof(int) * (count + 1))) != hipSuccess ||
      (e = b->d_counters.ensure(sizeof(int) * 8)) != hipSuccess ||
      (e = b->d_signs.ensure(sizeof(uint32_t) * 16 * (size_t)(b->ltot ? b->ltot : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  for (int i = 0; i < count; ++i)
    vout[i].signs = (decltype(vout[i].signs))(b->d_signs.as<uint32_t>() + (size_t)b->leaf_base[i] * 16);
  if ((e = hipMemcpyAsync(b->d_in.p, vin.data(), sizeof(GridView) * count, hipMemcpyHostToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(b->d_out_views.p, vout.data(), sizeof(GridView) * count, hipMemcpyHostToDevice,
                          ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(b->d_vel.p, vv.data(), sizeof(VelView) * count, hipMemcpyHostToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(b->d_outs.p, outs.data(), sizeof(NanoLeaf*) * count, hipMemcpyHostToDevice, ctx->stream)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(b->d_leaf_base.p, b->leaf_base.data(), sizeof(int) * (count + 1), hipMemcpyHostToDevice,
                          ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpy");
  int rc = batch_chunks(ctx, b.get());
  if (rc != VR_OK) return rc;
  *out = b.release();
  return VR_OK;
}

extern "C" int vr_batch_set_dt(vr_ctx* ctx, vr_batch* b, const double* dts) {
  if (!ctx || !b) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  // Captured graphs and mesh capacities hold for the old choice.
  batch_drop_graph(b);
  b->sized = false;
  if (!dts) {
    b->has_dts = false;
    return VR_OK;
  }
  hipError_t e;
  b->dts.assign(dts, dts + b->G);
  if ((e = b->d_dts.ensure(sizeof(double) * b->G)) != hipSuccess ||
      (e = hipMemcpyAsync(b->d_dts.p, dts, sizeof(double) * b->G, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpy");
  b->has_dts = true;
  return VR_OK;
}

extern "C" int vr_batch_set_tuning(vr_batch* b, int leaves_per_grab, int waves_per_cu) {
  if (!b || leaves_per_grab < 1 || waves_per_cu < 0 || waves_per_cu > 64) return VR_ERR_INVALID_ARGUMENT;
  b->lpg = leaves_per_grab;
  b->wpc = waves_per_cu;
  return VR_OK;
}

// Also writes the outputs' sign masks for `iso` (d_signs), which the
// batch's marching cubes classifies cells from.
static int batch_advect(vr_ctx* ctx, vr_batch* b, double dt, double t, int scheme, float iso) {
  if (b->chunk_lpg != b->lpg) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "chunk table out of date");
  double tf[3];
  // one field per grid; time factors from the first field's period (uniform fields use none)
  vr_stage_factors(b->vel[0]->period, t, dt, scheme, tf);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, ctx->stream, b->d_counters.as<int>());
  // wpc > 0: persistent, wpc waves per CU; wpc == 0: one block per chunk.
  // the batched kernels come in one- and four-wave forms
  const int nw = vr_advect_waves(b->ltot, b->cus) == 1 ? 1 : 4;
  int blocks = b->wpc > 0 ? b->cus * b->wpc / nw : 8 * b->max_range;
  if (blocks < 8) blocks = 8;
  const int persistent = b->wpc > 0 ? 1 : 0;
  const bool one = !persistent && b->lpg == 1;
  double bytes = 2.0 * VR_LEAF_BYTES * (double)b->ltot;
  {
    TimedLaunch tl(ctx, TC_ADVECT, 0.0, bytes, 0.0);
    const double* dts = b->has_dts ? b->d_dts.as<double>() : nullptr;
    uint32_t* signs = b->d_signs.as<uint32_t>();
    const int* leaf_base = b->d_leaf_base.as<int>();
    if (scheme == VR_SCHEME_RK3)
      launch_batch_s<3>(b->vk, nw, one, blocks, ctx->stream, b->d_in.as<GridView>(), b->d_vel.as<VelView>(),
                        b->d_outs.as<NanoLeaf*>(), b->d_chunks.as<int4>(), b->d_counters.as<int>(),
                        b->d_ranges.as<int>(), persistent, dts, dt, tf, signs, leaf_base, b->G, iso);
    else
      launch_batch_s<1>(b->vk, nw, one, blocks, ctx->stream, b->d_in.as<GridView>(), b->d_vel.as<VelView>(),
                        b->d_outs.as<NanoLeaf*>(), b->d_chunks.as<int4>(), b->d_counters.as<int>(),
                        b->d_ranges.as<int>(), persistent, dts, dt, tf, signs, leaf_base, b->G, iso);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "batch advect");
}

// Batched MC over the advected grids into the batch's mesh buffers (sizes
// fixed by the first, host-synchronised run).  The tile lists depend on the
// leaf sets only, which advection never changes: they are built once, by
// that first run, and every later step (and graph) starts at the count.
static int batch_mc(vr_ctx* ctx, vr_batch* b, float iso, bool sizing) {
  int rc;
  hipError_t e;
  if (!b->tiles_ready) {
    if (!sizing) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "batch tile lists not built");
    rc = mc_tiles_multi(ctx, b->mc, b->d_out_views.as<GridView>(), b->d_leaf_base.as<int>(), b->G, b->ltot);
    if (rc != VR_OK) return rc;
    int nt = 0;
    if ((e = hipMemcpyAsync(&nt, b->mc.totals.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "readback");
    b->tcap = nt;
    rc = mc_tiles_finish(ctx, b->mc, b->d_out_views.as<GridView>(), b->d_leaf_base.as<int>(), b->G, b->ltot, b->tcap);
    if (rc != VR_OK) return rc;
    b->tiles_ready = true;
  }
  long long waves = 1;
  rc = mc_count_multi(ctx, b->mc, b->d_out_views.as<GridView>(), b->ltot, b->tcap, iso, true, &waves);
  if (rc != VR_OK) return rc;
  double mesh_bytes = 0.0;  // known to the host after a sizing run only (timing is off in graphs)
  if (sizing) {
    // The count pass's mixed-cell lists get room for this step's demand (a
    // step of the same parameters has the same cells, so a graph replay
    // fits too); when they did not fit, the count runs again.
    unsigned long long share = 0;
    if ((e = hipMemcpyAsync(&share, b->mc.mctr.p, sizeof(share), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "readback");
    const unsigned long long want = share * (unsigned long long)waves;
    if (want > 0x7fffffffULL)
      return vr_fail(ctx, VR_ERR_CAPACITY, "batch needs %llu mixed-cell list entries (at most 2^31 - 1)", want);
    if (share > b->mc.mcell.bytes / sizeof(uint32_t) / (unsigned long long)waves) {
      batch_drop_graph(b);  // it reads the old lists
      if ((e = b->mc.mcell.ensure(sizeof(uint32_t) * (size_t)want)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
      rc = mc_count_multi(ctx, b->mc, b->d_out_views.as<GridView>(), b->ltot, b->tcap, iso, true);
      if (rc != VR_OK) return rc;
    }
    int th[3];
    if ((e = hipMemcpyAsync(th, b->mc.totals.p, sizeof(int) * 3, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "readback");
    mesh_bytes = 24.0 * th[1] + 12.0 * th[2];
    if (th[1] + 1 > b->vcap || th[2] + 1 > b->tricap) batch_drop_graph(b);  // it writes the old buffers
    if (th[1] + 1 > b->vcap) {
      if (b->d_verts) (void)hipFree(b->d_verts);
      b->d_verts = nullptr;
      if ((e = hipMalloc(&b->d_verts, sizeof(double) * 3 * (size_t)(th[1] + 1))) != hipSuccess)
        return vr_hip_fail(ctx, e, "hipMalloc");
      b->vcap = th[1] + 1;
    }
    if (th[2] + 1 > b->tricap) {
      if (b->d_tris) (void)hipFree(b->d_tris);
      b->d_tris = nullptr;
      if ((e = hipMalloc(&b->d_tris, sizeof(uint32_t) * 3 * (size_t)(th[2] + 1))) != hipSuccess)
        return vr_hip_fail(ctx, e, "hipMalloc");
      b->tricap = th[2] + 1;
    }
  }
  return mc_emit_multi(ctx, b->mc, b->d_out_views.as<GridView>(), b->G, b->tcap, iso, b->d_verts, b->d_tris,
                       (double)b->ltot, mesh_bytes, (long long)b->vcap, (long long)b->tricap, nullptr, true);
}

static int batch_read_ranges(vr_ctx* ctx, vr_batch* b) {
  b->ranges_host.assign((size_t)4 * b->G, 0);
  hipError_t e;
  if ((e = hipMemcpyAsync(b->ranges_host.data(), b->mc.ranges.p, sizeof(int) * 4 * b->G, hipMemcpyDeviceToHost,
                          ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "readback");
  return VR_OK;
}

extern "C" int vr_batch_step(vr_ctx* ctx, vr_batch* b, double dt, double t, int scheme, float iso, int mode) {
  if (!ctx || !b) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (scheme != VR_SCHEME_RK1 && scheme != VR_SCHEME_RK3)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "scheme must be 1 or 3");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  if (mode != VR_BATCH_LOOP) {
    const int rc = batch_chunks(ctx, b);
    if (rc != VR_OK) return rc;
  }
  if (mode == VR_BATCH_LOOP) {
    // The per-grid path: one advect launch and one marching-cubes pipeline per grid.
    if (b->loop_mesh.size() != (size_t)b->G) {
      b->loop_mesh.clear();
      for (int i = 0; i < b->G; ++i) b->loop_mesh.emplace_back(new vr_mesh());
    }
    for (int i = 0; i < b->G; ++i) {
      int rc = advect_into(ctx, b->in[i].get(), b->vel[i], b->has_dts ? b->dts[i] : dt, t, scheme, b->out[i].get());
      if (rc != VR_OK) return rc;
      if (b->out[i]->topo->count > 0) {
        MCScratch sc;
        rc = mc_run(ctx, b->out[i].get(), iso, &sc, b->loop_mesh[i].get());
        if (rc != VR_OK) return rc;
      }
    }
    return (e = hipStreamSynchronize(ctx->stream)) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "sync");
  }
  // Mesh capacities come from a host-synchronised run with the same
  // parameters (the tile count depends on the topology only, but vertex and
  // triangle counts depend on dt, t, the scheme, iso and the per-grid dts).
  if (b->sized && (b->s_dt != dt || b->s_t != t || b->s_scheme != scheme || b->s_iso != iso)) b->sized = false;
  if (!b->sized || mode == VR_BATCH_EAGER) {
    b->sized = false;
    int rc = batch_advect(ctx, b, dt, t, scheme, iso);
    if (rc == VR_OK) rc = batch_mc(ctx, b, iso, true);
    if (rc == VR_OK) rc = batch_read_ranges(ctx, b);
    if (rc != VR_OK) return rc;
    b->sized = true;
    b->s_dt = dt;
    b->s_t = t;
    b->s_scheme = scheme;
    b->s_iso = iso;
    if (mode == VR_BATCH_EAGER) return VR_OK;
  }
  // VR_BATCH_GRAPH: capture once per parameter set, then replay.
  if (!b->exec || b->g_dt != dt || b->g_t != t || b->g_scheme != scheme || b->g_iso != iso || b->g_lpg != b->lpg ||
      b->g_wpc != b->wpc) {
    batch_drop_graph(b);
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the captured graph
    if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
      ctx->timing = timing;
      return vr_hip_fail(ctx, e, "hipStreamBeginCapture");
    }
    int rc = batch_advect(ctx, b, dt, t, scheme, iso);
    if (rc == VR_OK) rc = batch_mc(ctx, b, iso, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = timing;
    if (rc != VR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return vr_hip_fail(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipGraphInstantiate");
    b->g_dt = dt;
    b->g_t = t;
    b->g_scheme = scheme;
    b->g_iso = iso;
    b->g_lpg = b->lpg;
    b->g_wpc = b->wpc;
  }
  {
    TimedLaunch tl(ctx, TC_GRAPH, 2.0 * VR_LEAF_BYTES, 0.0, (double)b->ltot);
    if ((e = hipGraphLaunch(b->exec, ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "hipGraphLaunch");
  }
  return VR_OK;
}

extern "C" int vr_batch_autotune(vr_ctx* ctx, vr_batch* b, double dt, int scheme, int* best_lpg, int* best_wpc) {
  if (!ctx || !b) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  hipEvent_t a, z;
  if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&z)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipEventCreate");
  const int lpgs[] = {1, 2, 4, 8, 16};
  const int wpcs[] = {0, 4, 8, 16, 32};  // 0: one block per chunk
  float best = 1e30f;
  int bl = b->lpg, bw = b->wpc;
  const bool timing = ctx->timing;
  ctx->timing = false;
  int rc = VR_OK;
  for (int lp : lpgs)
    for (int wp : wpcs) {
      b->lpg = lp;
      b->wpc = wp;
      rc = batch_chunks(ctx, b);
      if (rc == VR_OK) rc = batch_advect(ctx, b, dt, 0.0, scheme, 0.0f);  // warm
      (void)hipEventRecord(a, ctx->stream);
      for (int r = 0; r < 3 && rc == VR_OK; ++r) rc = batch_advect(ctx, b, dt, 0.0, scheme, 0.0f);
      (void)hipEventRecord(z, ctx->stream);
      if (rc != VR_OK || (e = hipEventSynchronize(z)) != hipSuccess) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, z);
      if (ms < best) {
        best = ms;
        bl = lp;
        bw = wp;
      }
    }
  ctx->timing = timing;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
  b->lpg = bl;
  b->wpc = bw;
  if (rc == VR_OK) rc = batch_chunks(ctx, b);
  if (best_lpg) *best_lpg = bl;
  if (best_wpc) *best_wpc = bw;
  return rc;
}

extern "C" int vr_batch_mesh_counts(vr_ctx* ctx, vr_batch* b, int i, int mode, uint64_t* nv, uint64_t* nt) {
  if (!b || i < 0 || i >= b->G) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad grid index");
  if (mode == VR_BATCH_LOOP) {
    if (b->loop_mesh.size() != (size_t)b->G) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "no loop step yet");
    if (nv) *nv = b->loop_mesh[i]->nv;
    if (nt) *nt = b->loop_mesh[i]->nt;
    return VR_OK;
  }
  if (!b->sized) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "no batched step yet");
  if (mode == VR_BATCH_GRAPH) {
    int rc = batch_read_ranges(ctx, b);
    if (rc != VR_OK) return rc;
  }
  const long long v0 = b->ranges_host[4 * i], t0 = b->ranges_host[4 * i + 2];
  const long long vn = b->ranges_host[4 * i + 1], tn = b->ranges_host[4 * i + 3];
  if (v0 < 0 || vn < 0 || t0 < 0 || tn < 0 || v0 + vn > b->vcap || t0 + tn > b->tricap)
    return vr_fail(ctx, VR_ERR_CAPACITY, "mesh of grid %d exceeds the batch's mesh buffers", i);
  if (nv) *nv = (uint64_t)vn;
  if (nt) *nt = (uint64_t)tn;
  return VR_OK;
}

extern "C" int vr_batch_mesh_copy(vr_ctx* ctx, vr_batch* b, int i, int mode, double* xyz, size_t vcap, uint32_t* tri,
                                  size_t tcap) {
  if (!ctx || !b || i < 0 || i >= b->G) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad grid index");
  if (mode == VR_BATCH_LOOP) {
    if (b->loop_mesh.size() != (size_t)b->G) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "no loop step yet");
    return vr_mesh_copy(ctx, b->loop_mesh[i].get(), xyz, vcap, tri, tcap);
  }
  uint64_t nv, nt;
  int rc = vr_batch_mesh_counts(ctx, b, i, mode, &nv, &nt);
  if (rc != VR_OK) return rc;
  if ((xyz && vcap < nv) || (tri && tcap < nt)) return vr_fail(ctx, VR_ERR_CAPACITY, "mesh buffers too small");
  hipError_t e;
  const size_t v0 = (size_t)b->ranges_host[4 * i], t0 = (size_t)b->ranges_host[4 * i + 2];
  if (xyz && nv &&
      (e = hipMemcpyAsync(xyz, b->d_verts + 3 * v0, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if (tri && nt &&
      (e = hipMemcpyAsync(tri, b->d_tris + 3 * t0, sizeof(uint32_t) * 3 * nt, hipMemcpyDeviceToHost, ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  return (e = hipStreamSynchronize(ctx->stream)) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "sync");
}

extern "C" int vr_batch_output(vr_batch* b, int i, const vr_grid** out) {
  if (!b || !out || i < 0 || i >= b->G) return VR_ERR_INVALID_ARGUMENT;
  *out = b->out[i].get();
  return VR_OK;
}

extern "C" int vr_batch_destroy(vr_ctx* ctx, vr_batch* b) {
  if (!b) return VR_ERR_INVALID_ARGUMENT;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete b;
  return VR_OK;
}
