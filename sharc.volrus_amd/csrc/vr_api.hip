// C ABI of the gfx950 narrow-band level-set library (include/volrus.h).
//
// Host side: NanoGrid validation and upload, the flattened root/internal/
// neighbour lookup arrays, device scratch management, and the sequencing of
// the kernels in vr_kernels.hip.  No torch, no CPU fallback: every compute
// call runs on the device or returns an error.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/volrus.h"
#include "mc_tables.h"
#include "vr_internal.h"

// ------------------------------------------------------------------ errors
int vr_fail(vr_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int vr_hip_fail(vr_ctx* ctx, hipError_t e, const char* where) {
  return vr_fail(ctx, e == hipErrorOutOfMemory ? VR_ERR_OUT_OF_MEMORY : VR_ERR_HIP, "%s: %s", where,
                 hipGetErrorString(e));
}

// ------------------------------------------------------------------ device buffers
DevBuf::~DevBuf() {
  if (p && owned) (void)hipFree(p);
}
hipError_t DevBuf::ensure(size_t n) {
  if (n <= bytes && p) return hipSuccess;
  if (!owned) return hipErrorOutOfMemory;  // borrowed memory is never reallocated
  if (p) {
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  const size_t want = n < 256 ? 256 : n;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess) bytes = want;
  return e;
}

Topo::~Topo() {
  if (d_count) (void)hipFree(d_count);
  if (d_origins) (void)hipFree(d_origins);
  if (d_roots) (void)hipFree(d_roots);
  if (d_child) (void)hipFree(d_child);
  if (d_nbr) (void)hipFree(d_nbr);
  if (d_count_alt) (void)hipFree(d_count_alt);
  if (d_nbr_alt) (void)hipFree(d_nbr_alt);
}

vr_grid::~vr_grid() {
  if (d_leaves) (void)hipFree(d_leaves);
}

// ------------------------------------------------------------------ MC tables
// Decodes the case table and derives the per-edge candidate-cell table used
// by the ownership rule, then copies both to constant memory once per device.
static std::mutex g_tables_mu;
static bool g_tables_done[64];

static hipError_t upload_tables_once(int device) {
  std::lock_guard<std::mutex> lk(g_tables_mu);
  if (device >= 0 && device < 64 && g_tables_done[device]) return hipSuccess;
  uint16_t em[256];
  uint8_t nt[256];
  int8_t tri[256][16];
  for (int c = 0; c < 256; ++c) {
    em[c] = (uint16_t)mc_edge_mask((uint32_t)c);
    const char* s = MC_TRI[c];
    const int len = (int)std::strlen(s);
    nt[c] = (uint8_t)(len / 3);
    for (int i = 0; i < 16; ++i) tri[c][i] = i < len ? (int8_t)mc_hex_digit(s[i]) : (int8_t)-1;
  }
  int8_t corner[8][3], edge[12][2], cand[12][4][4];
  for (int c = 0; c < 8; ++c)
    for (int a = 0; a < 3; ++a) corner[c][a] = (int8_t)MC_CORNER[c][a];
  for (int e = 0; e < 12; ++e) {
    edge[e][0] = (int8_t)MC_EDGE[e][0];
    edge[e][1] = (int8_t)MC_EDGE[e][1];
    const int* pa = MC_CORNER[MC_EDGE[e][0]];
    const int* pb = MC_CORNER[MC_EDGE[e][1]];
    int axis = 0;
    for (int a = 0; a < 3; ++a)
      if (pa[a] != pb[a]) axis = a;
    int lowp[3];
    for (int a = 0; a < 3; ++a) lowp[a] = pa[a] < pb[a] ? pa[a] : pb[a];
    const int u = (axis + 1) % 3, w = (axis + 2) % 3;
    int j = 0;
    for (int su = 0; su < 2; ++su)
      for (int sw = 0; sw < 2; ++sw) {
        int d[3] = {lowp[0], lowp[1], lowp[2]};
        d[u] -= su;
        d[w] -= sw;
        // the edge's corners in the candidate cell: lowp - d and lowp - d + axis
        int c0[3] = {lowp[0] - d[0], lowp[1] - d[1], lowp[2] - d[2]};
        int c1[3] = {c0[0], c0[1], c0[2]};
        c1[axis] += 1;
        int ep = -1;
        for (int f = 0; f < 12; ++f) {
          const int* qa = MC_CORNER[MC_EDGE[f][0]];
          const int* qb = MC_CORNER[MC_EDGE[f][1]];
          const bool same = (qa[0] == c0[0] && qa[1] == c0[1] && qa[2] == c0[2] && qb[0] == c1[0] && qb[1] == c1[1] &&
                             qb[2] == c1[2]) ||
                            (qb[0] == c0[0] && qb[1] == c0[1] && qb[2] == c0[2] && qa[0] == c1[0] && qa[1] == c1[1] &&
                             qa[2] == c1[2]);
          if (same) ep = f;
        }
        cand[e][j][0] = (int8_t)d[0];
        cand[e][j][1] = (int8_t)d[1];
        cand[e][j][2] = (int8_t)d[2];
        cand[e][j][3] = (int8_t)ep;
        ++j;
      }
  }
  hipError_t e = vr_upload_mc_tables(em, nt, &tri[0][0], &corner[0][0], &edge[0][0], &cand[0][0][0]);
  if (e == hipSuccess && device >= 0 && device < 64) g_tables_done[device] = true;
  return e;
}

// ------------------------------------------------------------------ context
extern "C" int vr_ctx_create(int device, vr_ctx** out) {
  if (!out) return VR_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return VR_ERR_HIP;
  if (device < 0 || device >= n) return VR_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return VR_ERR_HIP;
  std::unique_ptr<vr_ctx> c(new vr_ctx());
  c->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) c->cus = cus;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return VR_ERR_HIP;
  if (upload_tables_once(device) != hipSuccess || mc_noalloc_setup(device) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    return VR_ERR_HIP;
  }
  *out = c.release();
  return VR_OK;
}

vr_ctx::~vr_ctx() {
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" int vr_ctx_destroy(vr_ctx* ctx) {
  if (!ctx) return VR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return VR_OK;
}

extern "C" const char* vr_last_error(const vr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int vr_ctx_synchronize(vr_ctx* ctx) {
  if (!ctx) return VR_ERR_INVALID_ARGUMENT;
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "hipStreamSynchronize");
}

extern "C" void* vr_ctx_stream(vr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ------------------------------------------------------------------ topology
GridView make_view(const vr_grid* g) {
  GridView v;
  v.leaves = (decltype(v.leaves))g->d_leaves;
  v.count = (decltype(v.count))g->topo->d_count;
  v.origins = (decltype(v.origins))g->topo->d_origins;
  v.roots = (decltype(v.roots))g->topo->d_roots;
  v.child = (decltype(v.child))g->topo->d_child;
  v.nbr = (decltype(v.nbr))g->topo->d_nbr;
  v.signs = nullptr;
  v.root_count = (int)g->topo->roots.size();
  v.background = g->background;
  v.vs = g->voxel_size;
  v.inv_vs = 1.0 / g->voxel_size;
  v.diag = 1;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) {
      v.mat[4 * r + c] = g->mat[r][c];
      if (c < 3 && (c == r ? !(g->mat[r][c] > 0.0) : g->mat[r][c] != 0.0)) v.diag = 0;
    }
  return v;
}

// Allocates a topology for `cap` leaves over the given internal-node roots
// and builds child/origins/nbr from the leaves in `leaves` (count on device).
int build_topology(vr_ctx* ctx, vr_grid* g, std::shared_ptr<Topo> t) {
  hipError_t e;
  const int cap = t->cap < 1 ? 1 : t->cap;
  const size_t R = t->roots.size();
  if (!t->d_count && (e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMalloc(&t->d_origins, sizeof(int4) * cap)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMalloc(&t->d_roots, sizeof(int4) * (R ? R : 1))) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMalloc(&t->d_child, sizeof(int) * 4096 * (R ? R : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMalloc(&t->d_nbr, sizeof(int) * VR_NBR * (size_t)cap)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  if (R && (e = hipMemcpyAsync(t->d_roots, t->roots.data(), sizeof(int4) * R, hipMemcpyHostToDevice, ctx->stream)) !=
               hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if ((e = vr_fill_int(t->d_child, -1, 4096 * (R ? R : 1), ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "fill");
  if ((e = vr_build_child(g->d_leaves, t->d_count, cap, t->d_roots, (int)R, t->d_child, t->d_origins, ctx->stream)) !=
      hipSuccess)
    return vr_hip_fail(ctx, e, "build_child");
  g->topo = t;
  if ((e = vr_build_nbr(make_view(g), cap, t->d_nbr, ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "build_nbr");
  return VR_OK;
}

void copy_header_into(vr_grid* dst, const vr_grid* src);
static void copy_header(vr_grid* dst, const vr_grid* src) { copy_header_into(dst, src); }
void copy_header_into(vr_grid* dst, const vr_grid* src) {
  dst->grid_class = src->grid_class;
  dst->voxel_size = src->voxel_size;
  dst->background = src->background;
  std::memcpy(dst->mat, src->mat, sizeof(dst->mat));
  std::memcpy(dst->inv, src->inv, sizeof(dst->inv));
}

// Keeps the leaves of `src` (count = g's current count) flagged in `keep`,
// in order, as g's new leaf set, and rebuilds g's lookup arrays.  Mirrors
// remove_empty_leaves (src/tree/internal.rs:231-247) followed by the sorted
// leaf order of NanoGrid::from_grid.
int compact_into(vr_ctx* ctx, vr_grid* g, const NanoLeaf* src, const int* keep, bool has_empty_after, bool exact,
                 bool tail_clear, int* count_copy) {
  hipError_t e;
  const int n = g->topo->count;
  DevBuf& pos = ctx->scratch[SCR_POS];
  DevBuf& sums = ctx->scratch[SCR_SUMS];
  DevBuf& tot = ctx->scratch[SCR_NEWCOUNT];
  if ((e = pos.ensure(sizeof(int) * (size_t)(n + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
  if ((e = sums.ensure(sizeof(int) * (vr_scan_scratch_ints(n) + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if (g->d_leaves && g->d_leaves != src && g->topo.use_count() == 1) {
    // In place: the leaf buffer and lookup arrays are g's alone and large
    // enough (the count never grows), so reuse them.
    Topo* t = g->topo.get();
    if ((e = tot.ensure(sizeof(int) * 4)) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
    // With the host count an upper bound, the flags past the device count
    // are cleared so the scan sees only live leaves.
    // The compaction updates the lookup arrays too: the new count and
    // neighbour table go to the second buffers, which then become the
    // topology's.
    const size_t R = t->roots.size();
    if ((!t->d_count_alt && (e = hipMalloc(&t->d_count_alt, sizeof(int))) != hipSuccess) ||
        (!t->d_nbr_alt &&
         (e = hipMalloc(&t->d_nbr_alt, sizeof(int) * VR_NBR * (size_t)(t->cap < 1 ? 1 : t->cap))) != hipSuccess))
      return vr_hip_fail(ctx, e, "hipMalloc");
    if ((!exact && !tail_clear && (e = vr_zero_tail(const_cast<int*>(keep), t->d_count, n, ctx->stream)) != hipSuccess) ||
        (e = vr_exclusive_scan(keep, pos.as<int>(), (size_t)n, sums.as<int>(), tot.as<int>(), ctx->stream,
                               t->d_count)) != hipSuccess ||
        (e = vr_launch_compact_topo(src, t->d_count, n, keep, pos.as<int>(), tot.as<int>(), g->d_leaves, t->d_roots,
                                    (int)R, t->d_child, t->d_origins, t->d_nbr, t->d_nbr_alt, t->d_count_alt,
                                    count_copy, ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "compact");
    std::swap(t->d_count, t->d_count_alt);
    std::swap(t->d_nbr, t->d_nbr_alt);
    if (exact) {
      int cnt = 0;
      if ((e = hipMemcpyAsync(&cnt, tot.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
          (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return vr_hip_fail(ctx, e, "count readback");
      t->count = cnt;
    }
    t->has_empty = has_empty_after;
    return VR_OK;
  }
  std::shared_ptr<Topo> t(new Topo());
  t->cap = n;
  t->roots = g->topo->roots;
  if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = vr_exclusive_scan(keep, pos.as<int>(), (size_t)n, sums.as<int>(), t->d_count, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scan");
  NanoLeaf* dst = nullptr;
  if ((e = hipMalloc(&dst, sizeof(NanoLeaf) * (size_t)(n < 1 ? 1 : n))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = vr_launch_compact(src, g->topo->d_count, n, keep, pos.as<int>(), dst, ctx->stream)) != hipSuccess) {
    (void)hipFree(dst);
    return vr_hip_fail(ctx, e, "compact");
  }
  int cnt = 0;
  if ((e = hipMemcpyAsync(&cnt, t->d_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
    (void)hipFree(dst);
    return vr_hip_fail(ctx, e, "count readback");
  }
  t->count = cnt;
  t->has_empty = has_empty_after;
  if (count_copy && (e = vr_copy_int(t->d_count, count_copy, ctx->stream)) != hipSuccess) {
    (void)hipFree(dst);
    return vr_hip_fail(ctx, e, "count");
  }
  if (g->d_leaves) (void)hipFree(g->d_leaves);
  g->d_leaves = dst;
  g->cap = n;
  return build_topology(ctx, g, t);
}

// ------------------------------------------------------------------ upload / download
static bool origin_less(const int32_t* a, const int32_t* b) {
  return a[0] != b[0] ? a[0] < b[0] : (a[1] != b[1] ? a[1] < b[1] : a[2] < b[2]);
}

extern "C" int vr_grid_upload_nano(vr_ctx* ctx, const uint8_t* b, size_t len, vr_grid** out) {
  if (!ctx || !out || (!b && len)) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // src/nano/layout.rs:202-251
  if (len < NANO_HEADER_BYTES) return vr_fail(ctx, VR_ERR_INVALID_NANO, "Buffer too small for NanoHeader");
  if (std::memcmp(b, "NANO", 4) != 0) return vr_fail(ctx, VR_ERR_INVALID_NANO, "Invalid magic");
  uint32_t version, gclass, leaf_count;
  std::memcpy(&version, b + 4, 4);
  std::memcpy(&gclass, b + 8, 4);
  std::memcpy(&leaf_count, b + 28, 4);
  if (version != 1) return vr_fail(ctx, VR_ERR_INVALID_NANO, "Unsupported version: expected 1, got %u", version);
  const size_t expected = NANO_HEADER_BYTES + (size_t)VR_LEAF_BYTES * leaf_count;
  if (len < expected)
    return vr_fail(ctx, VR_ERR_INVALID_NANO, "Buffer too small: expected %zu bytes, got %zu", expected, len);
  const uint8_t* lb = b + NANO_HEADER_BYTES;
  std::vector<int4> roots;
  roots.reserve(16);
  bool has_empty = false;
  for (uint32_t i = 0; i < leaf_count; ++i) {
    int32_t o[3];
    std::memcpy(o, lb + (size_t)i * VR_LEAF_BYTES, 12);
    if (i > 0) {
      int32_t p[3];
      std::memcpy(p, lb + (size_t)(i - 1) * VR_LEAF_BYTES, 12);
      if (!origin_less(p, o)) return vr_fail(ctx, VR_ERR_INVALID_NANO, "Leaves not sorted at leaf %u", i);
    }
    if ((o[0] & 7) || (o[1] & 7) || (o[2] & 7))
      return vr_fail(ctx, VR_ERR_INVALID_NANO, "leaf %u origin (%d,%d,%d) is not a multiple of 8", i, o[0], o[1], o[2]);
    uint64_t m[8], any = 0;
    std::memcpy(m, lb + (size_t)i * VR_LEAF_BYTES + 16, 64);
    for (int w = 0; w < 8; ++w) any |= m[w];
    if (!any) has_empty = true;
    roots.push_back(make_int4(o[0] & ~127, o[1] & ~127, o[2] & ~127, 0));
  }
  std::sort(roots.begin(), roots.end(), [](const int4& a, const int4& c) {
    return a.x != c.x ? a.x < c.x : (a.y != c.y ? a.y < c.y : a.z < c.z);
  });
  roots.erase(std::unique(roots.begin(), roots.end(),
                          [](const int4& a, const int4& c) { return a.x == c.x && a.y == c.y && a.z == c.z; }),
              roots.end());

  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_grid> g(new vr_grid());
  std::memcpy(&g->voxel_size, b + 16, 8);
  std::memcpy(&g->background, b + 24, 4);
  std::memcpy(g->mat, b + 56, 128);
  std::memcpy(g->inv, b + 184, 128);
  g->grid_class = gclass;
  g->cap = (int)leaf_count;
  if ((e = hipMalloc(&g->d_leaves, sizeof(NanoLeaf) * (size_t)(leaf_count ? leaf_count : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc leaves");
  if (leaf_count &&
      (e = hipMemcpyAsync(g->d_leaves, lb, (size_t)VR_LEAF_BYTES * leaf_count, hipMemcpyHostToDevice, ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync leaves");
  std::shared_ptr<Topo> t(new Topo());
  t->cap = (int)leaf_count;
  t->count = (int)leaf_count;
  t->has_empty = has_empty;
  t->roots = std::move(roots);
  if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  const int cnt = (int)leaf_count;
  if ((e = hipMemcpyAsync(t->d_count, &cnt, sizeof(int), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  const int rc = build_topology(ctx, g.get(), t);
  if (rc != VR_OK) return rc;
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "upload");
  *out = g.release();
  return VR_OK;
}

extern "C" int vr_grid_download_nano(vr_ctx* ctx, const vr_grid* g, uint8_t* out, size_t cap, size_t* len) {
  if (!ctx || !g || !len) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  const int n = g->topo->count;
  const size_t need = NANO_HEADER_BYTES + (size_t)VR_LEAF_BYTES * n;
  *len = need;
  if (!out) return VR_OK;
  if (cap < need) return vr_fail(ctx, VR_ERR_CAPACITY, "need %zu bytes, have %zu", need, cap);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  if (n && (e = hipMemcpyAsync(out + NANO_HEADER_BYTES, g->d_leaves, (size_t)VR_LEAF_BYTES * n,
                               hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "download");
  // Header (src/nano/layout.rs:128-141); bbox over active voxels (:109-124).
  std::memset(out, 0, NANO_HEADER_BYTES);
  std::memcpy(out, "NANO", 4);
  const uint32_t version = 1, nc = (uint32_t)n;
  std::memcpy(out + 4, &version, 4);
  std::memcpy(out + 8, &g->grid_class, 4);
  std::memcpy(out + 16, &g->voxel_size, 8);
  std::memcpy(out + 24, &g->background, 4);
  std::memcpy(out + 28, &nc, 4);
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int i = 0; i < n; ++i) {
    const uint8_t* lp = out + NANO_HEADER_BYTES + (size_t)i * VR_LEAF_BYTES;
    int32_t o[3];
    uint64_t m[8];
    std::memcpy(o, lp, 12);
    std::memcpy(m, lp + 16, 64);
    uint64_t orw = 0;
    int xlo = 8, xhi = -1;
    for (int w = 0; w < 8; ++w)
      if (m[w]) {
        orw |= m[w];
        if (w < xlo) xlo = w;
        xhi = w;
      }
    if (!orw) continue;
    const int ylo = __builtin_ctzll(orw) >> 3, yhi = (63 - __builtin_clzll(orw)) >> 3;
    uint32_t zb = 0;
    for (int yy = 0; yy < 8; ++yy) zb |= (uint32_t)((orw >> (8 * yy)) & 0xFF);
    const int zlo = __builtin_ctz(zb), zhi = 31 - __builtin_clz(zb);
    mn[0] = std::min(mn[0], o[0] + xlo); mx[0] = std::max(mx[0], o[0] + xhi);
    mn[1] = std::min(mn[1], o[1] + ylo); mx[1] = std::max(mx[1], o[1] + yhi);
    mn[2] = std::min(mn[2], o[2] + zlo); mx[2] = std::max(mx[2], o[2] + zhi);
  }
  if (mn[0] == INT32_MAX) {
    for (int a = 0; a < 3; ++a) mn[a] = mx[a] = 0;
  }
  std::memcpy(out + 32, mn, 12);
  std::memcpy(out + 44, mx, 12);
  std::memcpy(out + 56, g->mat, 128);
  std::memcpy(out + 184, g->inv, 128);
  return VR_OK;
}

extern "C" int vr_grid_counts(vr_ctx* ctx, const vr_grid* g, uint64_t* leaf_count, uint64_t* active_count) {
  if (!ctx || !g) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  const int n = g->topo->count;
  if (leaf_count) *leaf_count = (uint64_t)n;
  if (!active_count) return VR_OK;
  hipError_t e;
  DevBuf& w = ctx->scratch[SCR_WORDS];
  DevBuf& pos = ctx->scratch[SCR_POS];
  DevBuf& sums = ctx->scratch[SCR_SUMS];
  DevBuf& tot = ctx->scratch[SCR_TOTAL];
  const size_t nw = (size_t)n * 8;
  if ((e = w.ensure(sizeof(int) * (nw + 1))) != hipSuccess || (e = pos.ensure(sizeof(int) * (nw + 1))) != hipSuccess ||
      (e = sums.ensure(sizeof(int) * (vr_scan_scratch_ints(nw) + 1))) != hipSuccess ||
      (e = tot.ensure(sizeof(int) * 4)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  // popcount of each mask word, summed by the scan's total
  DevBuf& masks = ctx->scratch[SCR_CROSS];
  (void)masks;
  if ((e = vr_launch_leaf_popcount(g->d_leaves, g->topo->d_count, n, w.as<int>(), ctx->stream)) != hipSuccess ||
      (e = vr_exclusive_scan(w.as<int>(), pos.as<int>(), nw, sums.as<int>(), tot.as<int>(), ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "count");
  int total = 0;
  if ((e = hipMemcpyAsync(&total, tot.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "count readback");
  *active_count = (uint64_t)total;
  return VR_OK;
}

extern "C" int vr_grid_destroy(vr_ctx* ctx, vr_grid* g) {
  if (!g) return VR_ERR_INVALID_ARGUMENT;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  delete g;
  return VR_OK;
}

// ------------------------------------------------------------------ velocity
extern "C" int vr_velocity_grids3(vr_ctx* ctx, const vr_grid* vx, const vr_grid* vy, const vr_grid* vz,
                                  vr_velocity** out) {
  if (!ctx || !vx || !vy || !vz || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  vr_velocity* v = new vr_velocity();
  v->kind = VK_GRIDS3;
  v->g[0] = vx;
  v->g[1] = vy;
  v->g[2] = vz;
  *out = v;
  return VR_OK;
}

extern "C" int vr_velocity_dense(vr_ctx* ctx, const float* uvw, int n, double h, vr_velocity** out) {
  if (!ctx || !uvw || !out || n <= 0 || !(h > 0.0)) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad dense field");
  std::unique_ptr<vr_velocity> v(new vr_velocity());
  v->kind = VK_DENSE;
  v->n = n;
  v->h = h;
  const size_t bytes = sizeof(float) * 3 * (size_t)n * n * n;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  if ((e = hipMalloc(&v->d_dense, bytes)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMemcpyAsync(v->d_dense, uvw, bytes, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpy");
  *out = v.release();
  return VR_OK;
}

extern "C" int vr_velocity_analytic(vr_ctx* ctx, int kind, const double params[6], vr_velocity** out) {
  if (!ctx || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (kind != VR_VEL_UNIFORM && kind != VR_VEL_RIGID && kind != VR_VEL_VORTEX)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "unknown analytic kind %d", kind);
  vr_velocity* v = new vr_velocity();
  v->kind = kind;
  for (int i = 0; i < 6; ++i) v->a[i] = params ? params[i] : 0.0;
  *out = v;
  return VR_OK;
}

extern "C" int vr_velocity_set_period(vr_velocity* v, double period) {
  if (!v) return VR_ERR_INVALID_ARGUMENT;
  v->period = period;
  return VR_OK;
}

vr_velocity::~vr_velocity() {
  if (d_dense) (void)hipFree(d_dense);
}

extern "C" int vr_velocity_destroy(vr_ctx* ctx, vr_velocity* v) {
  if (!v) return VR_ERR_INVALID_ARGUMENT;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete v;
  return VR_OK;
}

VelView make_vel_view(const vr_velocity* v) {
  VelView vv;
  std::memset(&vv, 0, sizeof(vv));
  vv.kind = v->kind;
  for (int i = 0; i < 6; ++i) vv.a[i] = v->a[i];
  if (v->kind == VK_GRIDS3)
    for (int c = 0; c < 3; ++c) vv.g[c] = make_view(v->g[c]);
  vv.dense = (decltype(vv.dense))v->d_dense;
  vv.n = v->n;
  vv.inv_h = v->kind == VK_DENSE ? 1.0 / v->h : 0.0;
  return vv;
}

double vr_time_factor(double period, double t) {
  if (!(period > 0.0)) return 1.0;
  double s, c;
  sincospi_det(t / period, &s, &c);
  return c;
}

void vr_stage_factors(double period, double t, double dt, int scheme, double tf[3]) {
  tf[0] = vr_time_factor(period, t + dt);
  tf[1] = scheme == VR_SCHEME_RK3 ? vr_time_factor(period, t) : 1.0;
  tf[2] = scheme == VR_SCHEME_RK3 ? vr_time_factor(period, t + 0.5 * dt) : 1.0;
}

// ------------------------------------------------------------------ advection
int advect_into(vr_ctx* ctx, const vr_grid* in, const vr_velocity* vel, double dt, double t, int scheme,
                vr_grid* o) {
  double tf[3];
  vr_stage_factors(vel->period, t, dt, scheme, tf);
  // Algorithmic bytes: read and write every leaf record once, plus the
  // velocity data (three leaf sets, or the dense array; 0 when analytic).
  double vel_bytes = 0.0;
  if (vel->kind == VK_GRIDS3)
    for (int c = 0; c < 3; ++c) vel_bytes += (double)VR_LEAF_BYTES * vel->g[c]->topo->count;
  if (vel->kind == VK_DENSE) vel_bytes += 12.0 * vel->n * vel->n * vel->n;
  hipError_t e;
  {
    TimedLaunch tl(ctx, TC_ADVECT, 2.0 * VR_LEAF_BYTES, vel_bytes, in->topo->count);
    e = vr_launch_advect(make_view(in), make_vel_view(vel), o->d_leaves, in->topo->cap, dt, tf, scheme,
                         vr_advect_waves(in->topo->count, ctx->cus), ctx->stream);
  }
  return e == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "advect");
}

extern "C" int vr_advect(vr_ctx* ctx, const vr_grid* in, const vr_velocity* vel, double dt, double t, int scheme,
                         vr_grid** out) {
  if (!ctx || !in || !vel || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (scheme != VR_SCHEME_RK1 && scheme != VR_SCHEME_RK3)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "scheme must be 1 or 3");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_grid> o(new vr_grid());
  copy_header(o.get(), in);
  o->topo = in->topo;
  o->cap = in->topo->count;
  if ((e = hipMalloc(&o->d_leaves, sizeof(NanoLeaf) * (size_t)(o->cap ? o->cap : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  int rc = advect_into(ctx, in, vel, dt, t, scheme, o.get());
  if (rc != VR_OK) return rc;
  if (in->topo->has_empty) {
    // The reference's output only holds leaves that received an active voxel.
    DevBuf& keep = ctx->scratch[SCR_KEEP];
    if ((e = keep.ensure(sizeof(int) * (size_t)(o->cap + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
    if ((e = vr_launch_nonempty(o->d_leaves, o->topo->d_count, o->cap, keep.as<int>(), ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "nonempty");
    NanoLeaf* src = o->d_leaves;
    o->d_leaves = nullptr;
    rc = compact_into(ctx, o.get(), src, keep.as<int>(), false);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(src);
    if (rc != VR_OK) return rc;
  }
  *out = o.release();
  return VR_OK;
}

// ------------------------------------------------------------------ band tools
int crossings_into(vr_ctx* ctx, const vr_grid* g, uint64_t** cross, int** any, int* gen) {
  hipError_t e;
  const int n = g->topo->count;
  DevBuf& cb = ctx->scratch[SCR_CROSS];
  DevBuf& ab = ctx->scratch[SCR_ANY];
  if ((e = cb.ensure(sizeof(uint64_t) * 8 * (size_t)(n + 1))) != hipSuccess ||
      (e = ab.ensure(sizeof(int) * 4)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  // The mark is the largest generation that found a crossing, so it needs
  // no reset between passes: only a new buffer or a wrapped counter clears it.
  if (ctx->cross_any != ab.p || ctx->cross_gen >= 0x7ffffff0) {
    if ((e = vr_fill_int(ab.as<int>(), 0, 1, ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "crossings");
    ctx->cross_any = ab.p;
    ctx->cross_gen = 0;
  }
  const int gn = ++ctx->cross_gen;
  {
    TimedLaunch tl(ctx, TC_CROSSINGS, (double)(VR_LEAF_BYTES + 64), 0.0, n);
    e = vr_launch_crossings(make_view(g), g->topo->cap, cb.as<uint64_t>(), ab.as<int>(), gn, ctx->stream);
  }
  if (e != hipSuccess) return vr_hip_fail(ctx, e, "crossings");
  *cross = cb.as<uint64_t>();
  *any = ab.as<int>();
  *gen = gn;
  return VR_OK;
}

extern "C" int vr_zero_crossings(vr_ctx* ctx, const vr_grid* g, int32_t* xyz, size_t cap, size_t* count) {
  if (!ctx || !g || !count) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  uint64_t* cross;
  int* any;
  int gen = 0;
  int rc = crossings_into(ctx, g, &cross, &any, &gen);
  if (rc != VR_OK) return rc;
  const int n = g->topo->count;
  const size_t nw = (size_t)n * 8;
  DevBuf& w = ctx->scratch[SCR_WORDS];
  DevBuf& pos = ctx->scratch[SCR_POS];
  DevBuf& sums = ctx->scratch[SCR_SUMS];
  DevBuf& tot = ctx->scratch[SCR_TOTAL];
  if ((e = w.ensure(sizeof(int) * (nw + 1))) != hipSuccess || (e = pos.ensure(sizeof(int) * (nw + 1))) != hipSuccess ||
      (e = sums.ensure(sizeof(int) * (vr_scan_scratch_ints(nw) + 1))) != hipSuccess ||
      (e = tot.ensure(sizeof(int) * 4)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if ((e = vr_launch_popcount_words(cross, g->topo->d_count, n, w.as<int>(), ctx->stream)) != hipSuccess ||
      (e = vr_exclusive_scan(w.as<int>(), pos.as<int>(), nw, sums.as<int>(), tot.as<int>(), ctx->stream)) !=
          hipSuccess)
    return vr_hip_fail(ctx, e, "crossing scan");
  int total = 0;
  if ((e = hipMemcpyAsync(&total, tot.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "readback");
  *count = (size_t)total;
  if (!xyz) return VR_OK;
  if (cap < (size_t)total) return vr_fail(ctx, VR_ERR_CAPACITY, "need %d coordinates", total);
  DevBuf& cb = ctx->scratch[SCR_COORDS];
  if ((e = cb.ensure(sizeof(int) * 3 * (size_t)(total + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
  if ((e = vr_launch_crossing_coords(g->topo->d_origins, g->topo->d_count, n, cross, pos.as<int>(), cb.as<int>(),
                                     ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "coords");
  if (total && (e = hipMemcpyAsync(xyz, cb.p, sizeof(int) * 3 * (size_t)total, hipMemcpyDeviceToHost, ctx->stream)) !=
                   hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "sync");
  return VR_OK;
}

extern "C" int vr_trim_band(vr_ctx* ctx, vr_grid* g, float half_width) {
  if (!ctx || !g) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  const int n = g->topo->count;
  const float vs = (float)g->voxel_size;
  const float threshold = half_width * vs;
  NanoLeaf* tmp = nullptr;
  DevBuf& keep = ctx->scratch[SCR_KEEP];
  if ((e = keep.ensure(sizeof(int) * (size_t)(n + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
  if ((e = hipMalloc(&tmp, sizeof(NanoLeaf) * (size_t)(n ? n : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  int rc = VR_OK;
  if ((e = vr_launch_trim(make_view(g), g->topo->cap, tmp, keep.as<int>(), threshold, ctx->stream)) != hipSuccess)
    rc = vr_hip_fail(ctx, e, "trim");
  if (rc == VR_OK) rc = compact_into(ctx, g, tmp, keep.as<int>(), false);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  return rc;
}

// Crossings, bounded-radius redistance into `tmp` (>= count leaves), then
// compaction back into g.
int rebuild_impl(vr_ctx* ctx, vr_grid* g, float half_width, NanoLeaf* tmp, bool exact, int* count_copy,
                 unsigned long long* active_out) {
  const double rr = std::ceil((double)half_width);
  if (!(rr <= 8.0)) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "rebuild supports ceil(half_width) <= 8");
  const int R = rr < 0.0 ? 0 : (int)rr;
  hipError_t e;
  const int n = g->topo->count;
  const float vs = (float)g->voxel_size;
  const float band = half_width * vs;
  uint64_t* cross;
  int* any;
  int gen = 0;
  int rc = crossings_into(ctx, g, &cross, &any, &gen);
  if (rc != VR_OK) return rc;
  DevBuf& keep = ctx->scratch[SCR_KEEP];
  if ((e = keep.ensure(sizeof(int) * (size_t)(n + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
  {
    TimedLaunch tl(ctx, TC_REBUILD, (double)(2 * VR_LEAF_BYTES + 64), 0.0, n);
    e = vr_launch_rebuild(make_view(g), g->topo->cap, cross, any, gen, tmp, keep.as<int>(), n, R, vs, band,
                          active_out, ctx->stream);
  }
  if (e != hipSuccess) return vr_hip_fail(ctx, e, "rebuild");
  // Leaves without active voxels survive only when the grid has no crossing
  // at all (every leaf is then kept as it was); the flag is read back only
  // when the grid can hold such leaves.
  int anyh = gen;
  if (g->topo->has_empty &&
      ((e = hipMemcpyAsync(&anyh, any, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
       (e = hipStreamSynchronize(ctx->stream)) != hipSuccess))
    return vr_hip_fail(ctx, e, "readback");
  return compact_into(ctx, g, tmp, keep.as<int>(), anyh == gen ? false : g->topo->has_empty, exact, true,
                      count_copy);
}

extern "C" int vr_rebuild_band(vr_ctx* ctx, vr_grid* g, float half_width) {
  if (!ctx || !g) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  DevBuf& tmp = ctx->scratch[SCR_TMP_LEAVES];
  if ((e = tmp.ensure(sizeof(NanoLeaf) * (size_t)(g->topo->count + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  return rebuild_impl(ctx, g, half_width, tmp.as<NanoLeaf>());
}

// ------------------------------------------------------------------ marching cubes
vr_mesh::~vr_mesh() {
  if (d_v) (void)hipFree(d_v);
  if (d_t) (void)hipFree(d_t);
}

extern "C" int vr_marching_cubes(vr_ctx* ctx, const vr_grid* g, float iso, vr_mesh** out) {
  if (!ctx || !g || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_mesh> m(new vr_mesh());
  const int n = g->topo->count;
  if (n == 0) {
    *out = m.release();
    return VR_OK;
  }
  MCScratch s;
  int rc = mc_run(ctx, g, iso, &s, m.get());
  if (rc != VR_OK) return rc;
  *out = m.release();
  return VR_OK;
}

extern "C" int vr_mesh_counts(vr_ctx* ctx, const vr_mesh* m, uint64_t* nv, uint64_t* nt) {
  if (!m) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null mesh");
  if (nv) *nv = m->nv;
  if (nt) *nt = m->nt;
  return VR_OK;
}

extern "C" int vr_mesh_copy(vr_ctx* ctx, const vr_mesh* m, double* xyz, size_t vcap, uint32_t* tri, size_t tcap) {
  if (!ctx || !m) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if ((xyz && vcap < m->nv) || (tri && tcap < m->nt))
    return vr_fail(ctx, VR_ERR_CAPACITY, "mesh needs %llu vertices, %llu triangles", (unsigned long long)m->nv,
                   (unsigned long long)m->nt);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  if (xyz && m->nv &&
      (e = hipMemcpyAsync(xyz, m->d_v, sizeof(double) * 3 * m->nv, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if (tri && m->nt &&
      (e = hipMemcpyAsync(tri, m->d_t, sizeof(uint32_t) * 3 * m->nt, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "sync");
  return VR_OK;
}

extern "C" int vr_mesh_destroy(vr_ctx* ctx, vr_mesh* m) {
  if (!m) return VR_ERR_INVALID_ARGUMENT;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete m;
  return VR_OK;
}
