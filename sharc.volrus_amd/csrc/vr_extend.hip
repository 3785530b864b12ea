// extend_velocity (src/tools/level_set_track.rs:127-169) on the device.
//
// Interface voxels: active with |phi| < vs; band voxels: active with
// vs <= |phi| < half_width * vs (f32, as the reference).  Each band voxel
// takes vx, vy, vz from the interface voxel at the least squared index
// distance; among equals the first in the reference's iteration order
// (min_by_key): internal-node origin, then leaf slot (lx, ly, lz), then voxel
// offset (x, y, z) -- the reference's order inside one 128^3 node, with
// nodes taken in ascending origin order where the reference's HashMap order
// is arbitrary.  set() activates the voxel, allocating its leaf when the
// velocity grid has none.
//
// The reference compares every band voxel with every interface voxel.  Here
// one block per leaf with band voxels stages the interface bits of its 27
// neighbouring leaves in LDS (24 x 24 rows of 24 bits) and searches the
// 17^3 window of Chebyshev radius 8 around each band voxel: per (x, y) row
// the nearest interface bit along z (the lower one on a tie: a lower z comes
// first in the order), then the least (d^2, order key) over the rows.  A
// voxel outside the window is at d^2 >= 81, so a window minimum below 81 is
// the global one, ties included; band voxels whose window holds none go to
// a brute-force pass over the list of all interface voxels.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "vr_internal.h"

// One wave per leaf: interface and band bits, per-leaf band flag, counts.
__global__ __launch_bounds__(256) void k_ev_classify(GridView g, int n, float vsf, float bandf, uint64_t* imask,
                                                     uint64_t* bmask, int* hasband, unsigned long long* counts) {
  const int lane = threadIdx.x & 63;
  for (int li = blockIdx.x * 4 + (threadIdx.x >> 6); li < n; li += gridDim.x * 4) {
    unsigned int ni = 0, nb = 0;
    uint64_t any = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const float v = g.leaves[li].values[w * 64 + lane];
      const bool act = (g.leaves[li].mask[w] >> lane) & 1ull;
      const float a = fabsf(v);
      const uint64_t ib = __ballot(act && a < vsf);
      const uint64_t bb = __ballot(act && a >= vsf && a < bandf);
      if (lane == 0) {
        imask[(size_t)li * 8 + w] = ib;
        bmask[(size_t)li * 8 + w] = bb;
      }
      ni += (unsigned)__popcll(ib);
      nb += (unsigned)__popcll(bb);
      any |= bb;
    }
    if (lane == 0) {
      hasband[li] = any ? 1 : 0;
      if (ni) atomicAdd(&counts[0], (unsigned long long)ni);
      if (nb) atomicAdd(&counts[1], (unsigned long long)nb);
    }
  }
}

// New leaf array of a velocity grid: old leaf src[j] copied, or (src[j] < 0)
// a new leaf at origin[j] holding the background, no voxel active.
__global__ __launch_bounds__(256) void k_ev_merge(const NanoLeaf* old, const int* src, const int4* origin, int n,
                                                  float bg, NanoLeaf* out) {
  for (int j = blockIdx.x; j < n; j += gridDim.x) {
    const int s = src[j];
    uint32_t* o = reinterpret_cast<uint32_t*>(out + j);
    if (s >= 0) {
      const uint32_t* p = reinterpret_cast<const uint32_t*>(old + s);
      for (int k = threadIdx.x; k < 532; k += blockDim.x) o[k] = p[k];
    } else {
      const int4 org = origin[j];
      for (int k = threadIdx.x; k < 532; k += blockDim.x) {
        uint32_t v;
        if (k == 0) v = (uint32_t)org.x;
        else if (k == 1) v = (uint32_t)org.y;
        else if (k == 2) v = (uint32_t)org.z;
        else if (k < 20) v = 0u;  // padding and mask
        else v = __float_as_uint(bg);
        o[k] = v;
      }
    }
  }
}

// Order key of a voxel within the window around leaf origin o: internal-node
// offset from o's node (-1, 0, +1 per axis), leaf slot, voxel offset.
__device__ __forceinline__ uint32_t ev_key(int x, int y, int z, int4 o) {
  const uint32_t rx = (uint32_t)((x >> 7) - (o.x >> 7) + 1), ry = (uint32_t)((y >> 7) - (o.y >> 7) + 1),
                 rz = (uint32_t)((z >> 7) - (o.z >> 7) + 1);
  uint32_t k = (rx * 4 + ry) * 4 + rz;
  k = ((k * 16 + (uint32_t)((x >> 3) & 15)) * 16 + (uint32_t)((y >> 3) & 15)) * 16 + (uint32_t)((z >> 3) & 15);
  return ((k * 8 + (uint32_t)(x & 7)) * 8 + (uint32_t)(y & 7)) * 8 + (uint32_t)(z & 7);
}

struct EVOut {
  NanoLeaf* leaves[3];
};

// Writes the three components at band voxel (x, y, z) from interface voxel
// (sx, sy, sz); bit `off` of the out leaves' masks is set by the caller.
__device__ __forceinline__ void ev_assign(const GridView* vel, const EVOut& out, const int* lout, int x, int y, int z,
                                          int sx, int sy, int sz) {
  const int off = voxel_offset(x, y, z);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float val = grid_get(vel[k], sx, sy, sz);
    const int li = lout ? lout[k] : find_leaf(vel[k], x, y, z);
    if (li >= 0) out.leaves[k][li].values[off] = val;
  }
}

// One block per sdf leaf with band voxels (blist).
__global__ __launch_bounds__(256) void k_ev_apply(GridView sdf, const uint64_t* imask, const uint64_t* bmask,
                                                  const int* blist, int nb, GridView v0, GridView v1, GridView v2,
                                                  EVOut out, int* fb_count, int4* fb_list) {
  __shared__ uint32_t rows[24 * 24];
  __shared__ unsigned long long lmask[3][8];
  __shared__ int lout[3];
  const GridView vel[3] = {v0, v1, v2};
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int li = blist[b];
    const int4 o = sdf.origins[li];
    for (int q = threadIdx.x; q < 24 * 24; q += blockDim.x) {
      const int xr = q / 24, yr = q % 24;
      uint32_t r = 0;
#pragma unroll
      for (int kz = 0; kz < 3; ++kz) {
        const int nl = sdf.nbr[(size_t)li * VR_NBR + (xr >> 3) * 9 + (yr >> 3) * 3 + kz];
        if (nl >= 0) r |= (uint32_t)((imask[(size_t)nl * 8 + (xr & 7)] >> ((yr & 7) * 8)) & 0xFFull) << (8 * kz);
      }
      rows[q] = r;
    }
    if (threadIdx.x < 24) lmask[threadIdx.x / 8][threadIdx.x % 8] = 0ull;
    if (threadIdx.x < 3) lout[threadIdx.x] = find_leaf(vel[threadIdx.x], o.x, o.y, o.z);
    __syncthreads();
    for (int v = threadIdx.x; v < 512; v += blockDim.x) {
      if (!((bmask[(size_t)li * 8 + (v >> 6)] >> (v & 63)) & 1ull)) continue;
      const int lx = v >> 6, ly = (v >> 3) & 7, lz = v & 7;
      const int X0 = lx + 8, Y0 = ly + 8, Z0 = lz + 8;
      uint64_t best = ~0ull;
      int bx = 0, by = 0, bz = 0;
      for (int dx = -8; dx <= 8; ++dx) {
        const uint32_t dx2 = (uint32_t)(dx * dx);
        if (dx2 > (uint32_t)(best >> 32)) continue;
        for (int dy = -8; dy <= 8; ++dy) {
          const uint32_t dxy = dx2 + (uint32_t)(dy * dy);
          if (dxy > (uint32_t)(best >> 32)) continue;
          const uint32_t r = rows[(X0 + dx) * 24 + Y0 + dy] & (0x1FFFFu << lz);  // z' in [lz, lz + 16]
          if (!r) continue;
          const uint32_t below = r & ((1u << Z0) - 1u), above = r >> Z0;
          const int db = below ? Z0 - (31 - __clz(below)) : 64;
          const int da = above ? __ffs(above) - 1 : 64;
          const int zr = db <= da ? Z0 - db : Z0 + da;  // nearer one; the lower on a tie
          const int dz = zr - Z0;
          const int gx = o.x - 8 + X0 + dx, gy = o.y - 8 + Y0 + dy, gz = o.z - 8 + zr;
          const uint64_t c = ((uint64_t)(dxy + (uint32_t)(dz * dz)) << 32) | ev_key(gx, gy, gz, o);
          if (c < best) {
            best = c;
            bx = gx;
            by = gy;
            bz = gz;
          }
        }
      }
      const int x = o.x + lx, y = o.y + ly, z = o.z + lz;
      if ((best >> 32) < 81ull) {
        ev_assign(vel, out, lout, x, y, z, bx, by, bz);
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicOr(&lmask[k][v >> 6], 1ull << (v & 63));
      } else {
        fb_list[atomicAdd(fb_count, 1)] = make_int4(x, y, z, 0);
      }
    }
    __syncthreads();
    if (threadIdx.x < 24) {
      const int k = threadIdx.x / 8, w = threadIdx.x % 8;
      const unsigned long long m = lmask[k][w];
      if (m && lout[k] >= 0) out.leaves[k][lout[k]].mask[w] |= m;
    }
    __syncthreads();
  }
}

// Lexicographic (d^2, internal-node origin, leaf slot, voxel offset).
__device__ __forceinline__ bool ev_before(long long d, int x, int y, int z, long long bd, int bx, int by, int bz) {
  if (d != bd) return d < bd;
  const int a[9] = {x >> 7, y >> 7, z >> 7, (x >> 3) & 15, (y >> 3) & 15, (z >> 3) & 15, x & 7, y & 7, z & 7};
  const int c[9] = {bx >> 7, by >> 7, bz >> 7, (bx >> 3) & 15, (by >> 3) & 15, (bz >> 3) & 15, bx & 7, by & 7, bz & 7};
  for (int i = 0; i < 9; ++i)
    if (a[i] != c[i]) return a[i] < c[i];
  return false;
}

// Band voxels whose window holds no interface voxel nearer than 9: all
// interface voxels (xyz triplets), one thread per band voxel.
__global__ __launch_bounds__(256) void k_ev_fallback(const int4* fb_list, const int* fb_count, const int* ixyz,
                                                     int ni, GridView v0, GridView v1, GridView v2, EVOut out) {
  const GridView vel[3] = {v0, v1, v2};
  const int nf = *fb_count;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const int4 p = fb_list[f];
    long long bd = -1;
    int bx = 0, by = 0, bz = 0;
    for (int i = 0; i < ni; ++i) {
      const int sx = ixyz[3 * i], sy = ixyz[3 * i + 1], sz = ixyz[3 * i + 2];
      const long long dx = (long long)p.x - sx, dy = (long long)p.y - sy, dz = (long long)p.z - sz;
      const long long d = dx * dx + dy * dy + dz * dz;
      if (bd < 0 || ev_before(d, sx, sy, sz, bd, bx, by, bz)) {
        bd = d;
        bx = sx;
        by = sy;
        bz = sz;
      }
    }
    if (bd < 0) continue;
    ev_assign(vel, out, nullptr, p.x, p.y, p.z, bx, by, bz);
    const int off = voxel_offset(p.x, p.y, p.z);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int li = find_leaf(vel[k], p.x, p.y, p.z);
      if (li >= 0) atomicOr(reinterpret_cast<unsigned long long*>(&out.leaves[k][li].mask[off >> 6]), 1ull << (off & 63));
    }
  }
}

static int ev_blocks(size_t n) { return n < 1 ? 1 : (n < 65536 ? (int)n : 65536); }

static bool int4_less(const int4& a, const int4& c) {
  return a.x != c.x ? a.x < c.x : (a.y != c.y ? a.y < c.y : a.z < c.z);
}
static bool int4_eq(const int4& a, const int4& c) { return a.x == c.x && a.y == c.y && a.z == c.z; }

// Gives velocity grid g a leaf at every origin of `need` (sorted) it lacks:
// a new leaf array with the old leaves copied and the new ones holding the
// background, and new lookup arrays (a topology shared with other grids
// stays theirs).
static int ev_grow(vr_ctx* ctx, vr_grid* g, const std::vector<int4>& need) {
  hipError_t e;
  const int n = g->topo->count;
  std::vector<int4> old((size_t)n);
  if (n && ((e = hipMemcpyAsync(old.data(), g->topo->d_origins, sizeof(int4) * n, hipMemcpyDeviceToHost, ctx->stream)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess))
    return vr_hip_fail(ctx, e, "origins");
  std::vector<int4> add;
  for (const int4& o : need)
    if (!std::binary_search(old.begin(), old.end(), o, int4_less)) add.push_back(o);
  if (add.empty()) return VR_OK;
  std::vector<int4> merged;
  std::vector<int> src;
  merged.reserve(old.size() + add.size());
  size_t a = 0, b = 0;
  while (a < old.size() || b < add.size()) {
    if (b >= add.size() || (a < old.size() && int4_less(old[a], add[b]))) {
      merged.push_back(old[a]);
      src.push_back((int)a++);
    } else {
      merged.push_back(add[b++]);
      src.push_back(-1);
    }
  }
  const int m = (int)merged.size();
  std::vector<int4> roots = g->topo->roots;
  for (const int4& o : add) roots.push_back(make_int4(o.x & ~127, o.y & ~127, o.z & ~127, 0));
  std::sort(roots.begin(), roots.end(), int4_less);
  roots.erase(std::unique(roots.begin(), roots.end(), int4_eq), roots.end());
  NanoLeaf* leaves = nullptr;
  DevBuf& dsrc = ctx->scratch[SCR_POS];
  DevBuf& dorg = ctx->scratch[SCR_COORDS];
  if ((e = hipMalloc(&leaves, sizeof(NanoLeaf) * (size_t)m)) != hipSuccess) return vr_hip_fail(ctx, e, "hipMalloc");
  if ((e = dsrc.ensure(sizeof(int) * m)) != hipSuccess || (e = dorg.ensure(sizeof(int4) * m)) != hipSuccess ||
      (e = hipMemcpyAsync(dsrc.p, src.data(), sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(dorg.p, merged.data(), sizeof(int4) * m, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) {
    (void)hipFree(leaves);
    return vr_hip_fail(ctx, e, "scratch");
  }
  hipLaunchKernelGGL(k_ev_merge, dim3(ev_blocks(m)), dim3(256), 0, ctx->stream, g->d_leaves, dsrc.as<int>(),
                     dorg.as<int4>(), m, g->background, leaves);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
    (void)hipFree(leaves);
    return vr_hip_fail(ctx, e, "merge");
  }
  std::shared_ptr<Topo> t(new Topo());
  t->cap = m;
  t->count = m;
  t->has_empty = g->topo->has_empty;  // every new leaf receives an active voxel
  t->roots = std::move(roots);
  if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess ||
      (e = hipMemcpyAsync(t->d_count, &t->count, sizeof(int), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) {
    (void)hipFree(leaves);
    return vr_hip_fail(ctx, e, "count");
  }
  if (g->d_leaves) (void)hipFree(g->d_leaves);
  g->d_leaves = leaves;
  g->cap = m;
  return build_topology(ctx, g, t);
}

extern "C" int vr_extend_velocity(vr_ctx* ctx, const vr_grid* sdf, vr_grid* vx, vr_grid* vy, vr_grid* vz,
                                  float half_width) {
  if (!ctx || !sdf || !vx || !vy || !vz) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (vx == vy || vy == vz || vx == vz) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "vx, vy, vz must be distinct grids");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  const int n = sdf->topo->count;
  if (n == 0) return VR_OK;  // no interface voxel: nothing changes
  const float vsf = (float)sdf->voxel_size;
  const float bandf = half_width * vsf;
  DevBuf& im = ctx->scratch[SCR_CROSS];
  DevBuf& bm = ctx->scratch[SCR_WORDS];
  DevBuf& hb = ctx->scratch[SCR_KEEP];
  DevBuf& cnt = ctx->scratch[SCR_ACTIVE];
  if ((e = im.ensure(sizeof(uint64_t) * 8 * (size_t)n)) != hipSuccess ||
      (e = bm.ensure(sizeof(uint64_t) * 8 * (size_t)n)) != hipSuccess ||
      (e = hb.ensure(sizeof(int) * (size_t)n)) != hipSuccess || (e = cnt.ensure(sizeof(unsigned long long) * 4)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if ((e = vr_fill_int(cnt.as<int>(), 0, 6, ctx->stream)) != hipSuccess) return vr_hip_fail(ctx, e, "fill");
  hipLaunchKernelGGL(k_ev_classify, dim3(ev_blocks(((size_t)n + 3) / 4)), dim3(256), 0, ctx->stream, make_view(sdf), n,
                     vsf, bandf, im.as<uint64_t>(), bm.as<uint64_t>(), hb.as<int>(), cnt.as<unsigned long long>());
  unsigned long long counts[2] = {0, 0};
  std::vector<int> hasband((size_t)n);
  std::vector<int4> org((size_t)n);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(counts, cnt.p, sizeof(counts), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(hasband.data(), hb.p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(org.data(), sdf->topo->d_origins, sizeof(int4) * n, hipMemcpyDeviceToHost, ctx->stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "classify");
  if (counts[0] == 0 || counts[1] == 0) return VR_OK;  // no interface, or no band voxel to fill
  std::vector<int> blist;
  std::vector<int4> need;
  for (int i = 0; i < n; ++i)
    if (hasband[i]) {
      blist.push_back(i);
      need.push_back(make_int4(org[i].x, org[i].y, org[i].z, 0));
    }
  vr_grid* vel[3] = {vx, vy, vz};
  for (int k = 0; k < 3; ++k) {
    const int rc = ev_grow(ctx, vel[k], need);
    if (rc != VR_OK) return rc;
  }
  const int nb = (int)blist.size();
  DevBuf& dbl = ctx->scratch[SCR_NEWCOUNT];
  DevBuf& fbl = ctx->scratch[SCR_TMP_LEAVES];
  if ((e = dbl.ensure(sizeof(int) * nb)) != hipSuccess ||
      (e = fbl.ensure(sizeof(int4) * (size_t)counts[1])) != hipSuccess ||
      (e = hipMemcpyAsync(dbl.p, blist.data(), sizeof(int) * nb, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  EVOut out{{vx->d_leaves, vy->d_leaves, vz->d_leaves}};
  int* fb_count = cnt.as<int>() + 4;
  hipLaunchKernelGGL(k_ev_apply, dim3(ev_blocks((size_t)nb)), dim3(256), 0, ctx->stream, make_view(sdf),
                     im.as<uint64_t>(), bm.as<uint64_t>(), dbl.as<int>(), nb, make_view(vx), make_view(vy),
                     make_view(vz), out, fb_count, fbl.as<int4>());
  int nf = 0;
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(&nf, fb_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "apply");
  if (nf > 0) {
    // All interface voxels as coordinates (popcount, scan, write), then the brute force.
    const size_t nw = (size_t)n * 8;
    DevBuf& w = ctx->scratch[SCR_MC_KEEP];
    DevBuf& pos = ctx->scratch[SCR_MC_POS];
    DevBuf& sums = ctx->scratch[SCR_SUMS];
    DevBuf& tot = ctx->scratch[SCR_TOTAL];
    DevBuf& xyz = ctx->scratch[SCR_MC_TILES];
    if ((e = w.ensure(sizeof(int) * (nw + 1))) != hipSuccess || (e = pos.ensure(sizeof(int) * (nw + 1))) != hipSuccess ||
        (e = sums.ensure(sizeof(int) * (vr_scan_scratch_ints(nw) + 1))) != hipSuccess ||
        (e = tot.ensure(sizeof(int) * 4)) != hipSuccess ||
        (e = xyz.ensure(sizeof(int) * 3 * (size_t)counts[0])) != hipSuccess)
      return vr_hip_fail(ctx, e, "scratch");
    if ((e = vr_launch_popcount_words(im.as<uint64_t>(), sdf->topo->d_count, n, w.as<int>(), ctx->stream)) !=
            hipSuccess ||
        (e = vr_exclusive_scan(w.as<int>(), pos.as<int>(), nw, sums.as<int>(), tot.as<int>(), ctx->stream)) !=
            hipSuccess ||
        (e = vr_launch_crossing_coords(sdf->topo->d_origins, sdf->topo->d_count, n, im.as<uint64_t>(), pos.as<int>(),
                                       xyz.as<int>(), ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "interface list");
    hipLaunchKernelGGL(k_ev_fallback, dim3(ev_blocks(((size_t)nf + 255) / 256)), dim3(256), 0, ctx->stream,
                       fbl.as<int4>(), fb_count, xyz.as<int>(), (int)counts[0], make_view(vx), make_view(vy),
                       make_view(vz), out);
    if ((e = hipGetLastError()) != hipSuccess) return vr_hip_fail(ctx, e, "fallback");
  }
  return (e = hipStreamSynchronize(ctx->stream)) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "extend_velocity");
}
