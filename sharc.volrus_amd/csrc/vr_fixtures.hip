// Synthetic level-set generation on the device.
//
// A scene is a left fold of primitives, ((p0 op1 p1) op2 p2) ..., each
// rasterised like make_level_set_sphere (src/tools/sphere.rs:17-59: the f64
// distance is cast to f32 and stored iff |value| < bg, bg = (f32)(hw * vs))
// and combined like csg_combine (src/tools/csg.rs:32-78: a voxel is kept if
// either operand is active there or the result differs from the background
// by more than 1e-12).  The host lists candidate leaves (leaf boxes that can
// reach the band, by a Lipschitz bound); one wave per candidate leaf then
// evaluates its 512 voxels, and empty leaves are compacted away.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vr_internal.h"

struct PrimDev {
  int kind;  // VR_PRIM_*
  int op;    // VR_CSG_* (ignored for the first primitive)
  double p[8];
  int lo[3], hi[3];  // index box visited by the rasteriser loop
};

__host__ __device__ inline double noise_lattice(long long i, long long j, long long k, unsigned long long seed) {
  unsigned long long h = seed ^ ((unsigned long long)i * 0x9E3779B97F4A7C15ull) ^
                         ((unsigned long long)j * 0xC2B2AE3D27D4EB4Full) ^ ((unsigned long long)k * 0x165667B19E3779F9ull);
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Four octaves of trilinear value noise with a cubic fade (frequency f0 2^o,
// weight 2^-o), scaled so |noise| <= amp.
__host__ __device__ inline double value_noise4(double wx, double wy, double wz, double amp, double f0,
                                               unsigned long long seed) {
  double sum = 0.0, wsum = 0.0, freq = f0, wt = 1.0;
  const double w[3] = {wx, wy, wz};
  for (int o = 0; o < 4; ++o) {
    double fr[3];
    long long b[3];
    for (int a = 0; a < 3; ++a) {
      const double g = w[a] * freq;
      const double fl = floor(g);
      b[a] = (long long)fl;
      const double t = g - fl;
      fr[a] = t * t * (3.0 - 2.0 * t);
    }
    double acc = 0.0;
    for (int q = 0; q < 8; ++q) {
      const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
      const double wq = (dx ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) * (dz ? fr[2] : 1.0 - fr[2]);
      acc += wq * noise_lattice(b[0] + dx, b[1] + dy, b[2] + dz, seed + (unsigned long long)o);
    }
    sum += wt * acc;
    wsum += wt;
    freq *= 2.0;
    wt *= 0.5;
  }
  return amp * (sum / wsum);
}

__host__ __device__ inline double prim_sdf(const PrimDev& pr, double wx, double wy, double wz) {
  if (pr.kind == VR_PRIM_SPHERE || pr.kind == VR_PRIM_NOISY_SPHERE) {
    const double dx = wx - pr.p[0], dy = wy - pr.p[1], dz = wz - pr.p[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz) - pr.p[3];
    if (pr.kind == VR_PRIM_NOISY_SPHERE) d += value_noise4(wx, wy, wz, pr.p[4], pr.p[5], (unsigned long long)pr.p[6]);
    return d;
  }
  const double w[3] = {wx, wy, wz};
  double outside = 0.0, inside = -1e300;
  for (int a = 0; a < 3; ++a) {
    const double c = 0.5 * (pr.p[a] + pr.p[a + 3]), hlf = 0.5 * (pr.p[a + 3] - pr.p[a]);
    const double q = fabs(w[a] - c) - hlf;
    const double m = q > 0.0 ? q : 0.0;
    outside += m * m;
    inside = q > inside ? q : inside;
  }
  return sqrt(outside) + (inside < 0.0 ? inside : 0.0);
}

__device__ __forceinline__ float csg_min_f(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return b < a ? b : a;
}
__device__ __forceinline__ float csg_max_f(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return b > a ? b : a;
}

__global__ __launch_bounds__(64) void k_gen_leaves(const int4* cand, int n, const PrimDev* prims, int np, float bg,
                                                   double vs, NanoLeaf* out, int* keep) {
  const int lane = threadIdx.x;
  const int y = lane >> 3, z = lane & 7;
  int li_end, li_step;
  for (int li = xcd_first(n, &li_end, &li_step); li < li_end; li += li_step) {
    const int4 o = cand[li];
    NanoLeaf& O = out[li];
    if (lane < 3) O.origin[lane] = lane == 0 ? o.x : (lane == 1 ? o.y : o.z);
    if (lane == 3) O.pad = 0;
    bool any = false;
    for (int x = 0; x < 8; ++x) {
      const int i = o.x + x, j = o.y + y, k = o.z + z;
      const double wx = (double)i * vs, wy = (double)j * vs, wz = (double)k * vs;
      float v = bg;
      bool act = false;
      for (int q = 0; q < np; ++q) {
        const PrimDev& pr = prims[q];
        float pv = bg;
        bool pa = false;
        if (i >= pr.lo[0] && i <= pr.hi[0] && j >= pr.lo[1] && j <= pr.hi[1] && k >= pr.lo[2] && k <= pr.hi[2]) {
          const float s = (float)prim_sdf(pr, wx, wy, wz);
          if (fabsf(s) < bg) {
            pv = s;
            pa = true;
          }
        }
        if (q == 0) {
          v = pv;
          act = pa;
          continue;
        }
        const float comb = pr.op == VR_CSG_UNION ? csg_min_f(v, pv)
                           : pr.op == VR_CSG_INTERSECTION ? csg_max_f(v, pv)
                                                          : csg_max_f(v, -pv);
        act = act || pa || fabsf(comb - bg) > 1e-12f;
        v = act ? comb : bg;
      }
      O.values[x * 64 + lane] = v;
      const uint64_t word = __ballot(act);
      if (lane == 0) O.mask[x] = word;
      any |= word != 0;
    }
    if (lane == 0) keep[li] = any ? 1 : 0;
  }
}

// Index box of the rasteriser loop for one primitive (sphere.rs:32-40 shape).
static void prim_box(const vr_primitive& p, double vs, double band, int lo[3], int hi[3]) {
  double a[3], b[3];
  if (p.kind == VR_PRIM_BOX) {
    for (int k = 0; k < 3; ++k) {
      a[k] = p.p[k] - band;
      b[k] = p.p[k + 3] + band;
    }
  } else {
    const double ext = p.p[3] + band + (p.kind == VR_PRIM_NOISY_SPHERE ? p.p[4] : 0.0);
    for (int k = 0; k < 3; ++k) {
      a[k] = p.p[k] - ext;
      b[k] = p.p[k] + ext;
    }
  }
  const double inv = 1.0 / vs;
  for (int k = 0; k < 3; ++k) {
    lo[k] = (int)std::floor(a[k] * inv);
    hi[k] = (int)std::ceil(b[k] * inv);
  }
}

extern "C" int vr_grid_from_primitives(vr_ctx* ctx, const vr_primitive* prims, int count, double voxel_size,
                                       double half_width, vr_grid** out) {
  if (!ctx || !prims || count <= 0 || !out || !(voxel_size > 0.0) || !(half_width > 0.0))
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad primitive list");
  *out = nullptr;
  const double band = half_width * voxel_size;
  const float bg = (float)band;
  std::vector<PrimDev> pd((size_t)count);
  std::vector<long long> keys;
  const double halfdiag = std::sqrt(3.0) * 3.5 * voxel_size;
  for (int q = 0; q < count; ++q) {
    const vr_primitive& p = prims[q];
    if (p.kind < VR_PRIM_SPHERE || p.kind > VR_PRIM_NOISY_SPHERE || (q > 0 && (p.op < 0 || p.op > 2)))
      return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "primitive %d: bad kind/op", q);
    PrimDev& d = pd[q];
    d.kind = p.kind;
    d.op = p.op;
    for (int k = 0; k < 8; ++k) d.p[k] = p.p[k];
    prim_box(p, voxel_size, band, d.lo, d.hi);
    for (int k = 0; k < 3; ++k)
      if (d.hi[k] - d.lo[k] > (1 << 22)) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "primitive %d too large", q);
    // Candidate leaves: leaf box can hold a voxel with |sdf| < band.
    PrimDev core = d;
    if (core.kind == VR_PRIM_NOISY_SPHERE) core.kind = VR_PRIM_SPHERE;
    const double slack = band + halfdiag + (p.kind == VR_PRIM_NOISY_SPHERE ? p.p[4] : 0.0) + 1e-3 * voxel_size;
    for (int lx = d.lo[0] & ~7; lx <= d.hi[0]; lx += 8)
      for (int ly = d.lo[1] & ~7; ly <= d.hi[1]; ly += 8)
        for (int lz = d.lo[2] & ~7; lz <= d.hi[2]; lz += 8) {
          const double s = prim_sdf(core, (lx + 3.5) * voxel_size, (ly + 3.5) * voxel_size, (lz + 3.5) * voxel_size);
          if (std::fabs(s) <= slack) {
            const long long key = (((long long)(lx >> 3) + (1 << 20)) << 42) | (((long long)(ly >> 3) + (1 << 20)) << 21) |
                                  ((long long)(lz >> 3) + (1 << 20));
            keys.push_back(key);
          }
        }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int n = (int)keys.size();
  std::vector<int4> cand((size_t)n);
  std::vector<int4> roots;
  for (int i = 0; i < n; ++i) {
    const long long k = keys[i];
    const int x = (int)(((k >> 42) & ((1 << 21) - 1)) - (1 << 20)) * 8;
    const int y = (int)(((k >> 21) & ((1 << 21) - 1)) - (1 << 20)) * 8;
    const int z = (int)((k & ((1 << 21) - 1)) - (1 << 20)) * 8;
    cand[i] = make_int4(x, y, z, 0);
    roots.push_back(make_int4(x & ~127, y & ~127, z & ~127, 0));
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
  g->grid_class = 1;  // LevelSet
  g->voxel_size = voxel_size;
  g->background = bg;
  // Grid::set_transform(VoxelTransform::uniform(vs)) (src/grid.rs:180-199):
  // the inverse translation is -0.0 * (1/vs) = -0.0.
  std::memset(g->mat, 0, sizeof(g->mat));
  std::memset(g->inv, 0, sizeof(g->inv));
  for (int r = 0; r < 3; ++r) {
    g->mat[r][r] = voxel_size;
    g->inv[r][r] = 1.0 / voxel_size;
    g->inv[r][3] = -0.0 * (1.0 / voxel_size);
  }
  g->mat[3][3] = g->inv[3][3] = 1.0;
  DevBuf& cb = ctx->scratch[SCR_COORDS];
  DevBuf& pb = ctx->scratch[SCR_WORDS];
  DevBuf& keep = ctx->scratch[SCR_KEEP];
  NanoLeaf* tmp = nullptr;
  if ((e = cb.ensure(sizeof(int4) * (size_t)(n + 1))) != hipSuccess ||
      (e = pb.ensure(sizeof(PrimDev) * (size_t)count)) != hipSuccess ||
      (e = keep.ensure(sizeof(int) * (size_t)(n + 1))) != hipSuccess ||
      (e = hipMalloc(&tmp, sizeof(NanoLeaf) * (size_t)(n ? n : 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  int rc = VR_OK;
  if (n && ((e = hipMemcpyAsync(cb.p, cand.data(), sizeof(int4) * n, hipMemcpyHostToDevice, ctx->stream)) !=
                hipSuccess ||
            (e = hipMemcpyAsync(pb.p, pd.data(), sizeof(PrimDev) * count, hipMemcpyHostToDevice, ctx->stream)) !=
                hipSuccess))
    rc = vr_hip_fail(ctx, e, "hipMemcpyAsync");
  if (rc == VR_OK && n) {
    const int blocks = n < 65536 ? n : 65536;
    hipLaunchKernelGGL(k_gen_leaves, dim3(blocks), dim3(64), 0, ctx->stream, cb.as<int4>(), n, pb.as<PrimDev>(), count,
                       bg, voxel_size, tmp, keep.as<int>());
    if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "gen");
  }
  if (rc == VR_OK) {
    std::shared_ptr<Topo> t(new Topo());
    t->cap = n;
    t->count = n;
    t->roots = roots;
    if ((e = hipMalloc(&t->d_count, sizeof(int))) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipMalloc");
    if (rc == VR_OK && (e = hipMemcpyAsync(t->d_count, &n, sizeof(int), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "hipMemcpyAsync");
    if (rc == VR_OK) {
      g->topo = t;
      g->cap = n;
      rc = compact_into(ctx, g.get(), tmp, keep.as<int>(), false);
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  if (rc != VR_OK) return rc;
  *out = g.release();
  return VR_OK;
}
