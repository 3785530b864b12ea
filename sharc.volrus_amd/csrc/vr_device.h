// Device-side data layout and helpers shared by the gfx950 kernels.
//
// HBM layout of one level set (a "grid"):
//   leaves   NanoLeaf[cap]     the NanoGrid leaf record, 2128 B each, sorted by
//                              origin (src/nano/layout.rs:41-52): origin at 0,
//                              active mask (8 x u64) at 16, values (512 f32,
//                              x*64 + y*8 + z) at 80.
//   count    int               live leaf count (device resident, so launches
//                              never need it on the host).
//   origins  int4[cap]         compact copy of the leaf origins for searches.
//   roots    int4[R]           sorted origins of the 128^3 internal nodes.
//   child    int[R * 4096]     leaf index per internal slot (-1: no leaf),
//                              slot = (lx << 8) | (ly << 4) | lz
//                              (src/tree/internal.rs:88-95).
//   nbr      int[cap * 27]     index of the leaf at origin + 8 * (dx, dy, dz),
//                              entry (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VR_LEAF_BYTES 2128
#define VR_NBR 27

struct __attribute__((aligned(8))) NanoLeaf {
  int32_t origin[3];
  int32_t pad;
  uint64_t mask[8];
  float values[512];
};
static_assert(sizeof(NanoLeaf) == VR_LEAF_BYTES, "NanoLeaf must be the 2128-byte NanoGrid record");

// Read-only view of a grid passed to kernels by value.
// Grid arrays always live in global memory.  Views are also read from
// device memory (batches, multi-grid marching cubes), where a plain pointer
// would compile to FLAT accesses (which also wait on the LDS counter); in the
// device pass the pointers are typed global.  The host pass sees plain
// pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define VR_GLOBAL __attribute__((address_space(1)))
#else
#define VR_GLOBAL
#endif

struct GridView {
  const VR_GLOBAL NanoLeaf* leaves;
  const VR_GLOBAL int* count;
  const VR_GLOBAL int4* origins;
  const VR_GLOBAL int4* roots;
  const VR_GLOBAL int* child;
  const VR_GLOBAL int* nbr;
  // Optional sign masks, 16 words per leaf: bit (offset & 31) of word
  // offset >> 5 = value at `offset` < the iso value they were made for
  // (written by the batched advection for its marching cubes; else null).
  const VR_GLOBAL uint32_t* signs;
  int root_count;
  float background;
  int diag;         // mat has zero off-diagonal terms and a positive diagonal
  double vs;        // voxel_size of the uniform transform
  double inv_vs;    // 1.0 / voxel_size (grid.rs:114-121)
  double mat[12];   // forward affine rows 0..2 (grid.rs:248-257)
};

// ------------------------------------------------------------------ lookups
__device__ __forceinline__ bool coord_less(int ax, int ay, int az, int bx, int by, int bz) {
  return ax != bx ? ax < bx : (ay != by ? ay < by : az < bz);
}

// Leaf index of the leaf containing voxel (x, y, z), or -1.
__device__ __forceinline__ int find_leaf(const GridView& g, int x, int y, int z) {
  const int kx = x & ~127, ky = y & ~127, kz = z & ~127;
  int lo = 0, hi = g.root_count;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int4 r = g.roots[mid];
    if (coord_less(r.x, r.y, r.z, kx, ky, kz))
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= g.root_count) return -1;
  const int4 r = g.roots[lo];
  if (r.x != kx || r.y != ky || r.z != kz) return -1;
  const int slot = (((x - kx) >> 3) << 8) | (((y - ky) >> 3) << 4) | ((z - kz) >> 3);
  return g.child[(size_t)lo * 4096 + slot];
}

__device__ __forceinline__ int voxel_offset(int x, int y, int z) { return ((x & 7) << 6) | ((y & 7) << 3) | (z & 7); }

// Grid::get (src/grid.rs:484-486 -> tree lookup): stored value inside an
// allocated leaf (active or not), background elsewhere.
__device__ __forceinline__ float grid_get(const GridView& g, int x, int y, int z) {
  const int li = find_leaf(g, x, y, z);
  return li < 0 ? g.background : g.leaves[li].values[voxel_offset(x, y, z)];
}

__device__ __forceinline__ bool grid_is_active(const GridView& g, int x, int y, int z) {
  const int li = find_leaf(g, x, y, z);
  if (li < 0) return false;
  const int off = voxel_offset(x, y, z);
  return (g.leaves[li].mask[off >> 6] >> (off & 63)) & 1ull;
}

// Rust `f64 as i32`: saturating, NaN -> 0.  The hardware conversion
// v_cvt_i32_f64 does exactly that (out-of-range values and infinities
// saturate, NaN gives 0), in one instruction instead of a clamp, a
// conversion and a NaN select; C++'s (int) cast leaves out-of-range values
// undefined, so the instruction is named here.
__device__ __forceinline__ int f64_to_i32(double v) {
  int r;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// The 7-lerp blend of src/tools/interpolation.rs:39-47, in f32, in that order.
__device__ __forceinline__ float trilinear_blend(const float c[8], float fx, float fy, float fz) {
  // c index = dx | dy << 1 | dz << 2
  const float c00 = c[0] * (1.0f - fx) + c[1] * fx;
  const float c01 = c[4] * (1.0f - fx) + c[5] * fx;
  const float c10 = c[2] * (1.0f - fx) + c[3] * fx;
  const float c11 = c[6] * (1.0f - fx) + c[7] * fx;
  const float c0 = c00 * (1.0f - fy) + c10 * fy;
  const float c1 = c01 * (1.0f - fy) + c11 * fy;
  return c0 * (1.0f - fz) + c1 * fz;
}

// sample_trilinear (src/tools/interpolation.rs:11-48) through the tree.
__device__ __forceinline__ float sample_grid(const GridView& g, double wx, double wy, double wz) {
  const double ix = wx * g.inv_vs, iy = wy * g.inv_vs, iz = wz * g.inv_vs;
  const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
  const float fx = (float)(ix - flx), fy = (float)(iy - fly), fz = (float)(iz - flz);
  const int i = f64_to_i32(flx), j = f64_to_i32(fly), k = f64_to_i32(flz);
  float c[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) c[q] = grid_get(g, i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
  return trilinear_blend(c, fx, fy, fz);
}

// AffineMap::index_to_world (src/grid.rs:248-257), same association order.
__device__ __forceinline__ void index_to_world(const double m[12], int x, int y, int z, double w[3]) {
  const double fx = (double)x, fy = (double)y, fz = (double)z;
#pragma unroll
  for (int r = 0; r < 3; ++r) w[r] = m[4 * r] * fx + m[4 * r + 1] * fy + m[4 * r + 2] * fz + m[4 * r + 3];
}

// The same for a view; with zero off-diagonal terms each product with them
// is +-0, and adding +-0 leaves the partial sum m_rr * x_r unchanged (it is
// never -0: m_rr > 0 and the integer converts to +0, never -0), so the row
// is m_rr * x_r + m_r3, bit for bit.
__device__ __forceinline__ void view_index_to_world(const GridView& g, int x, int y, int z, double w[3]) {
  if (g.diag) {
    w[0] = g.mat[0] * (double)x + g.mat[3];
    w[1] = g.mat[5] * (double)y + g.mat[7];
    w[2] = g.mat[10] * (double)z + g.mat[11];
  } else {
    index_to_world(g.mat, x, y, z, w);
  }
}

// Stages the 10^3 brick (local coordinates -1..8) around the leaf whose 27
// neighbour indices are nb[] (nb[13] is the leaf itself) into LDS; missing
// leaves read as the background.  One wave, lane in [0, 64).
__device__ __forceinline__ void stage_brick_dev(const GridView& in, const int* nb, float* brick, int lane) {
  for (int i = lane; i < 1000; i += 64) {
    const int bx = i / 100 - 1, by = (i / 10) % 10 - 1, bz = i % 10 - 1;
    const int dx = bx < 0 ? -1 : (bx > 7 ? 1 : 0), dy = by < 0 ? -1 : (by > 7 ? 1 : 0),
              dz = bz < 0 ? -1 : (bz > 7 ? 1 : 0);
    const int src = nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
    brick[i] = src < 0 ? in.background : in.leaves[src].values[voxel_offset(bx, by, bz)];
  }
}

// ------------------------------------------------------------------ sinpi / cospi
// Exact reduction of x to f in [-1/4, 1/4] and the Taylor series of sin(pi f)
// and cos(pi f) to f^15 / f^14 (truncation below 4e-16 and 1e-15), Horner with
// explicit fused multiply-adds (each a single IEEE rounding), so host and
// device produce the same bits.
__host__ __device__ inline void sincospi_det(double x, double* so, double* co) {
  const double n = rint(2.0 * x);
  const double f = x - 0.5 * n;
  const double f2 = f * f;
  double s = -2.1915353447830217e-05;
  s = fma(s, f2, 0.00046630280576761255);
  s = fma(s, f2, -0.0073704309457143504);
  s = fma(s, f2, 0.08214588661112823);
  s = fma(s, f2, -0.5992645293207921);
  s = fma(s, f2, 2.5501640398773455);
  s = fma(s, f2, -5.16771278004997);
  s = fma(s, f2, 3.141592653589793);
  s = s * f;
  double c = -0.0001046381049248457;
  c = fma(c, f2, 0.0019295743094039231);
  c = fma(c, f2, -0.02580689139001406);
  c = fma(c, f2, 0.2353306303588932);
  c = fma(c, f2, -1.3352627688545895);
  c = fma(c, f2, 4.0587121264167685);
  c = fma(c, f2, -4.934802200544679);
  c = fma(c, f2, 1.0);
  const long long q = ((long long)n) & 3;
  if (q == 0) { *so = s; *co = c; }
  else if (q == 1) { *so = c; *co = -s; }
  else if (q == 2) { *so = -s; *co = -c; }
  else { *so = -c; *co = s; }
}

// ------------------------------------------------------------------ velocity
enum VelKind { VK_GRIDS3 = 0, VK_DENSE = 1, VK_UNIFORM = 2, VK_RIGID = 3, VK_VORTEX = 4 };

struct VelView {
  int kind;
  double a[6];           // analytic parameters
  GridView g[3];         // VK_GRIDS3: vx, vy, vz
  const VR_GLOBAL float* dense;  // VK_DENSE: 3 * n^3, component-major, (i * n + j) * n + k
  int n;
  double inv_h;          // 1.0 / node spacing
};

__device__ __forceinline__ float dense_at(const float* u, int n, int i, int j, int k) {
  if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n || (unsigned)k >= (unsigned)n) return 0.0f;
  return u[((size_t)i * n + j) * n + k];
}

// Velocity at world point p (double), before the time factor.
template <int VK>
__device__ __forceinline__ void vel_sample(const VelView& v, const double p[3], double out[3]) {
  if constexpr (VK == VK_GRIDS3) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = (double)sample_grid(v.g[c], p[0], p[1], p[2]);
  } else if constexpr (VK == VK_DENSE) {
    const double ix = p[0] * v.inv_h, iy = p[1] * v.inv_h, iz = p[2] * v.inv_h;
    const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
    const float fx = (float)(ix - flx), fy = (float)(iy - fly), fz = (float)(iz - flz);
    const int i = f64_to_i32(flx), j = f64_to_i32(fly), k = f64_to_i32(flz);
    const size_t nn = (size_t)v.n * v.n * v.n;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const float* u = v.dense + comp * nn;
      float c[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) c[q] = dense_at(u, v.n, i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
      out[comp] = (double)trilinear_blend(c, fx, fy, fz);
    }
  } else if constexpr (VK == VK_UNIFORM) {
    out[0] = v.a[0]; out[1] = v.a[1]; out[2] = v.a[2];
  } else if constexpr (VK == VK_RIGID) {
    const double rx = p[0] - v.a[3], ry = p[1] - v.a[4], rz = p[2] - v.a[5];
    out[0] = v.a[1] * rz - v.a[2] * ry;
    out[1] = v.a[2] * rx - v.a[0] * rz;
    out[2] = v.a[0] * ry - v.a[1] * rx;
  } else {
    double sx, cx, sy, cy;
    sincospi_det(p[0], &sx, &cx);
    sincospi_det(p[1], &sy, &cy);
    out[0] = (sx * sx) * (2.0 * sy * cy);
    out[1] = -((sy * sy) * (2.0 * sx * cx));
    out[2] = 0.0;
  }
}

// Departure point of the back-trace from x (RK1: the reference's
// x - dt v(x), src/tools/advection.rs:40-48; RK3: SSP Shu-Osher on the
// backward characteristic).  tf = stage time factors.
template <int VK, int SCHEME>
__device__ __forceinline__ void departure_from(const VelView& v, const double x[3], const double u0[3], double dt,
                                              const double tf[3], double out[3]) {
  double u[3] = {u0[0], u0[1], u0[2]};
  if constexpr (SCHEME == 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = x[a] - dt * (u[a] * tf[0]);
  } else {
    double y1[3], y2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) y1[a] = x[a] - dt * (u[a] * tf[0]);
    vel_sample<VK>(v, y1, u);
#pragma unroll
    for (int a = 0; a < 3; ++a) y2[a] = 0.75 * x[a] + 0.25 * (y1[a] - dt * (u[a] * tf[1]));
    vel_sample<VK>(v, y2, u);
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = (1.0 / 3.0) * x[a] + (2.0 / 3.0) * (y2[a] - dt * (u[a] * tf[2]));
  }
}

// The same, with the velocity at x sampled here (u0 = v(x)).
template <int VK, int SCHEME>
__device__ __forceinline__ void departure(const VelView& v, const double x[3], double dt, const double tf[3],
                                         double out[3]) {
  double u0[3];
  vel_sample<VK>(v, x, u0);
  departure_from<VK, SCHEME>(v, x, u0, dt, tf, out);
}

// ------------------------------------------------------------------ XCD-aware item order
// Items [0, n) are cut into 8 contiguous ranges, one per XCD: workgroups are
// dispatched to the 8 XCDs round-robin (workgroup b on XCD b % 8), so items
// next to each other (leaves whose halos overlap, tiles sharing leaves) are
// processed under one XCD's L2 instead of being fetched by several.  Block b
// starts at the returned item and steps by *stride up to *end.  Grids of
// fewer than 8 blocks use the plain order.
__device__ __forceinline__ int xcd_first(int n, int* end, int* stride) {
  const int B = (int)gridDim.x, b = (int)blockIdx.x;
  if (B < 8) {
    *end = n;
    *stride = B;
    return b;
  }
  const int x = b & 7;
  *stride = (B - x + 7) >> 3;  // blocks on this XCD
  *end = (int)((long long)n * (x + 1) / 8);
  return (int)((long long)n * x / 8) + (b >> 3);
}

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Inclusive prefix sum over the 64 lanes of a wave; every lane must be
// active.  DPP row shifts scan each row of 16 lanes (lanes whose source is
// outside the row add 0), then row broadcasts carry lane 15 into row 1 and
// row 3, and lane 31 into rows 2 and 3.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}
