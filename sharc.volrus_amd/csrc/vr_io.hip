// .vol files (src/io/vol_format.rs:53-147) to and from device grids.
//
// A .vol file is a 32-byte header and one 2124-byte record per leaf
// (origin, 8 x u64 active mask, 512 f32 values), in any leaf order.  The
// device keeps NanoGrid's 2128-byte leaf record (origin, 4 bytes of padding,
// mask, values) sorted by origin, so reading is: check the header and the
// record count on the host, order the records by their leaf origin (origin
// & ~7, the leaf read_vol allocates; a later record of the same leaf
// replaces an earlier one, as read_vol's overwrite does), copy the file's
// records to HBM as they are and repack them there into sorted leaves, one
// block per leaf.  Writing repacks sorted leaves into records on the device
// and copies them out once; the file then holds the leaves in ascending
// origin order (the reference writes its HashMap order; every order reads
// back as the same grid).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "vr_internal.h"

#define VOL_HEADER 32
#define VOL_RECORD 2124          // bytes per leaf record
#define VOL_RECORD_WORDS 531     // 32-bit words per record
#define NANO_LEAF_WORDS 532      // 32-bit words per NanoLeaf

// Leaf i of the sorted grid from record perm[i] of the file (words of
// VOL_RECORD_WORDS, 4-byte aligned: the header and records are multiples of
// 4 bytes).  Leaf word 3 is NanoLeaf's padding; the origin is the aligned
// one the host computed.
__global__ __launch_bounds__(256) void k_vol_unpack(const uint32_t* rec, const int* perm, const int4* origin, int n,
                                                    NanoLeaf* leaves) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t* r = rec + (size_t)perm[i] * VOL_RECORD_WORDS;
    uint32_t* l = reinterpret_cast<uint32_t*>(leaves + i);
    const int4 o = origin[i];
    for (int k = threadIdx.x; k < NANO_LEAF_WORDS; k += blockDim.x) {
      uint32_t v;
      if (k == 0) v = (uint32_t)o.x;
      else if (k == 1) v = (uint32_t)o.y;
      else if (k == 2) v = (uint32_t)o.z;
      else if (k == 3) v = 0u;
      else v = r[k - 1];
      l[k] = v;
    }
  }
}

// Record i of the file from leaf i (NanoLeaf word 3, the padding, dropped).
__global__ __launch_bounds__(256) void k_vol_pack(const NanoLeaf* leaves, const int* count, uint32_t* rec) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t* l = reinterpret_cast<const uint32_t*>(leaves + i);
    uint32_t* r = rec + (size_t)i * VOL_RECORD_WORDS;
    for (int k = threadIdx.x; k < VOL_RECORD_WORDS; k += blockDim.x) r[k] = l[k < 3 ? k : k + 1];
  }
}

static int vol_blocks(int n) { return n < 1 ? 1 : (n < 65536 ? n : 65536); }

extern "C" int vr_grid_upload_vol(vr_ctx* ctx, const uint8_t* b, size_t len, vr_grid** out) {
  if (!ctx || !out || (!b && len)) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // vol_format.rs:90-113 (read_exact: a short buffer is UnexpectedEof)
  if (len < VOL_HEADER) return vr_fail(ctx, VR_ERR_INVALID_VOL, "failed to fill whole buffer");
  if (std::memcmp(b, "VOLR", 4) != 0) return vr_fail(ctx, VR_ERR_INVALID_VOL, "not a .vol file (bad magic)");
  uint32_t version, cls, n;
  double vs;
  float bg;
  std::memcpy(&version, b + 4, 4);
  if (version != 1) return vr_fail(ctx, VR_ERR_INVALID_VOL, "unsupported .vol version: %u", version);
  std::memcpy(&cls, b + 8, 4);
  if (cls > 2) return vr_fail(ctx, VR_ERR_INVALID_VOL, "unknown grid class: %u", cls);
  std::memcpy(&vs, b + 12, 8);
  std::memcpy(&bg, b + 20, 4);
  std::memcpy(&n, b + 24, 4);
  if ((len - VOL_HEADER) / VOL_RECORD < (size_t)n) return vr_fail(ctx, VR_ERR_INVALID_VOL, "failed to fill whole buffer");
  if (n > 0x7fffffffu / 2) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "too many leaves");
  // Leaf order: by aligned origin, the last record of a leaf wins.
  const uint8_t* rb = b + VOL_HEADER;
  std::vector<int> idx(n);
  std::vector<int4> org(n);
  for (uint32_t i = 0; i < n; ++i) {
    int32_t o[3];
    std::memcpy(o, rb + (size_t)i * VOL_RECORD, 12);
    org[i] = make_int4(o[0] & ~7, o[1] & ~7, o[2] & ~7, 0);
    idx[i] = (int)i;
  }
  auto less = [&](int a, int c) {
    const int4 &p = org[a], &q = org[c];
    return p.x != q.x ? p.x < q.x : (p.y != q.y ? p.y < q.y : p.z < q.z);
  };
  std::stable_sort(idx.begin(), idx.end(), less);
  std::vector<int> perm;
  std::vector<int4> origins;
  perm.reserve(n);
  origins.reserve(n);
  for (size_t k = 0; k < idx.size(); ++k) {
    if (k + 1 < idx.size() && !less(idx[k], idx[k + 1])) continue;  // a later record of the same leaf follows
    perm.push_back(idx[k]);
    origins.push_back(org[idx[k]]);
  }
  const int L = (int)perm.size();
  bool has_empty = false;
  std::vector<int4> roots;
  for (int i = 0; i < L; ++i) {
    uint64_t m[8], any = 0;
    std::memcpy(m, rb + (size_t)perm[i] * VOL_RECORD + 12, 64);
    for (int w = 0; w < 8; ++w) any |= m[w];
    if (!any) has_empty = true;
    const int4 r = make_int4(origins[i].x & ~127, origins[i].y & ~127, origins[i].z & ~127, 0);
    if (roots.empty() || roots.back().x != r.x || roots.back().y != r.y || roots.back().z != r.z) roots.push_back(r);
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
  // Grid::new(background, voxel_size) + set_grid_class (vol_format.rs:115-116):
  // uniform scale, zero translation.
  g->voxel_size = vs;
  g->background = bg;
  g->grid_class = cls;
  std::memset(g->mat, 0, sizeof(g->mat));
  std::memset(g->inv, 0, sizeof(g->inv));
  for (int a = 0; a < 3; ++a) {
    g->mat[a][a] = vs;
    g->inv[a][a] = 1.0 / vs;
  }
  g->mat[3][3] = g->inv[3][3] = 1.0;
  g->cap = L;
  if ((e = hipMalloc(&g->d_leaves, sizeof(NanoLeaf) * (size_t)(L ? L : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc leaves");
  if (L) {
    DevBuf& rec = ctx->scratch[SCR_TMP_LEAVES];
    DevBuf& dperm = ctx->scratch[SCR_POS];
    DevBuf& dorg = ctx->scratch[SCR_COORDS];
    if ((e = rec.ensure((size_t)n * VOL_RECORD)) != hipSuccess || (e = dperm.ensure(sizeof(int) * L)) != hipSuccess ||
        (e = dorg.ensure(sizeof(int4) * L)) != hipSuccess)
      return vr_hip_fail(ctx, e, "scratch");
    if ((e = hipMemcpyAsync(rec.p, rb, (size_t)n * VOL_RECORD, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(dperm.p, perm.data(), sizeof(int) * L, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(dorg.p, origins.data(), sizeof(int4) * L, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess)
      return vr_hip_fail(ctx, e, "hipMemcpyAsync");
    hipLaunchKernelGGL(k_vol_unpack, dim3(vol_blocks(L)), dim3(256), 0, ctx->stream, rec.as<uint32_t>(),
                       dperm.as<int>(), dorg.as<int4>(), L, g->d_leaves);
    if ((e = hipGetLastError()) != hipSuccess) return vr_hip_fail(ctx, e, "vol unpack");
  }
  std::shared_ptr<Topo> t(new Topo());
  t->cap = L;
  t->count = L;
  t->has_empty = has_empty;
  t->roots = std::move(roots);
  if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = hipMemcpyAsync(t->d_count, &t->count, sizeof(int), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  const int rc = build_topology(ctx, g.get(), t);
  if (rc != VR_OK) return rc;
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "upload");
  *out = g.release();
  return VR_OK;
}

extern "C" int vr_grid_download_vol(vr_ctx* ctx, const vr_grid* g, uint8_t* out, size_t cap, size_t* len) {
  if (!ctx || !g || !len) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  const int n = g->topo->count;
  const size_t need = VOL_HEADER + (size_t)VOL_RECORD * n;
  *len = need;
  if (!out) return VR_OK;
  if (cap < need) return vr_fail(ctx, VR_ERR_CAPACITY, "need %zu bytes, have %zu", need, cap);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  if (n) {
    DevBuf& rec = ctx->scratch[SCR_TMP_LEAVES];
    if ((e = rec.ensure((size_t)n * VOL_RECORD)) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
    hipLaunchKernelGGL(k_vol_pack, dim3(vol_blocks(n)), dim3(256), 0, ctx->stream, g->d_leaves, g->topo->d_count,
                       rec.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(out + VOL_HEADER, rec.p, (size_t)n * VOL_RECORD, hipMemcpyDeviceToHost, ctx->stream)) !=
            hipSuccess)
      return vr_hip_fail(ctx, e, "vol pack");
  }
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "download");
  // vol_format.rs:55-62
  std::memcpy(out, "VOLR", 4);
  const uint32_t version = 1, cls = g->grid_class, nc = (uint32_t)n;
  std::memcpy(out + 4, &version, 4);
  std::memcpy(out + 8, &cls, 4);
  std::memcpy(out + 12, &g->voxel_size, 8);
  std::memcpy(out + 20, &g->background, 4);
  std::memcpy(out + 24, &nc, 4);
  std::memset(out + 28, 0, 4);
  return VR_OK;
}
