// Per-voxel and per-leaf advection device code shared by the single-grid
// kernel (vr_kernels.hip) and the batched persistent kernel (vr_batch.hip).
#pragma once

#include "vr_device.h"

// New value of the active voxel (ox + x, oy + y, oz + z): back-trace, then
// trilinear sampling of the input (src/tools/advection.rs:39-53).  The
// departure cell's corners come from the LDS brick (leaf plus one-voxel halo,
// 10^3) when they all lie in it; otherwise each goes through the
// root/internal lookup (any dt), one corner at a time to keep that rare path
// small in registers.
// vtab (used when tab: single-vortex field on a diagonal transform): per
// local x and y, sin and cos of pi times the voxel row's world coordinate,
// so the velocity at the voxel itself (the first stage) takes no sincospi.
template <int VK, int SCHEME>
__device__ __forceinline__ float advect_voxel(const GridView& in, const VelView& vel, const float* brick, int ox,
                                              int oy, int oz, int x, int y, int z, double dt, const double tf[3],
                                              bool tab, const double* vtab) {
  double w[3], p[3];
  view_index_to_world(in, ox + x, oy + y, oz + z, w);
  if constexpr (VK == VK_VORTEX) {
    if (tab) {
      // vel_sample<VK_VORTEX> at w from the row tables: the same values
      const double sx = vtab[x], cx = vtab[8 + x], sy = vtab[16 + y], cy = vtab[24 + y];
      const double u0[3] = {(sx * sx) * (2.0 * sy * cy), -((sy * sy) * (2.0 * sx * cx)), 0.0};
      departure_from<VK, SCHEME>(vel, w, u0, dt, tf, p);
    } else {
      departure<VK, SCHEME>(vel, w, dt, tf, p);
    }
  } else {
    departure<VK, SCHEME>(vel, w, dt, tf, p);
  }
  const double ix = p[0] * in.inv_vs, iy = p[1] * in.inv_vs, iz = p[2] * in.inv_vs;
  const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
  const float fx = (float)(ix - flx), fy = (float)(iy - fly), fz = (float)(iz - flz);
  const int i = f64_to_i32(flx), j = f64_to_i32(fly), k = f64_to_i32(flz);
  const int bi = i - ox, bj = j - oy, bk = k - oz;
  float c[8];
  if (bi >= -1 && bi <= 7 && bj >= -1 && bj <= 7 && bk >= -1 && bk <= 7) {
    const int b0 = (bi + 1) * 100 + (bj + 1) * 10 + (bk + 1);
#pragma unroll
    for (int q = 0; q < 8; ++q) c[q] = brick[b0 + (q & 1) * 100 + ((q >> 1) & 1) * 10 + ((q >> 2) & 1)];
  } else {
#pragma unroll 1
    for (int q = 0; q < 8; ++q) c[q] = grid_get(in, i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
  }
  return trilinear_blend(c, fx, fy, fz);
}

// Stages the 10^3 brick of leaf li (leaf plus one-voxel halo) with NT
// threads in two parts: the leaf's own 512 values straight from its record
// (no dependence on the neighbour list, so they load while the list does),
// then the 488 halo values through nb[] once the caller has loaded it into
// LDS and synchronised.
template <int NT>
__device__ __forceinline__ void stage_interior(const GridView& in, int li, float* brick, int tid) {
  const float4* own = reinterpret_cast<const float4*>(in.leaves[li].values);
  for (int q = tid; q < 128; q += NT) {
    const float4 v = own[q];
    const int x = q >> 4, y = (q >> 1) & 7, z = (q & 1) * 4;
    float* b = brick + (x + 1) * 100 + (y + 1) * 10 + (z + 1);
    b[0] = v.x;
    b[1] = v.y;
    b[2] = v.z;
    b[3] = v.w;
  }
}

// Halo: the 1000 - 512 brick positions with some local coordinate -1 or 8,
// as a table built at compile time: brick index | neighbour slot (0..26) <<
// 10 | voxel offset in that neighbour << 15.
struct AdvectHalo {
  uint32_t e[488];
};
constexpr AdvectHalo make_advect_halo() {
  AdvectHalo t{};
  int h = 0;
  for (int bx = -1; bx <= 8; ++bx)
    for (int by = -1; by <= 8; ++by)
      for (int bz = -1; bz <= 8; ++bz) {
        if (bx >= 0 && bx <= 7 && by >= 0 && by <= 7 && bz >= 0 && bz <= 7) continue;
        const int dx = bx < 0 ? -1 : (bx > 7 ? 1 : 0), dy = by < 0 ? -1 : (by > 7 ? 1 : 0),
                  dz = bz < 0 ? -1 : (bz > 7 ? 1 : 0);
        const uint32_t b = (uint32_t)((bx + 1) * 100 + (by + 1) * 10 + (bz + 1));
        const uint32_t d = (uint32_t)((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1));
        const uint32_t off = (uint32_t)(((bx & 7) << 6) | ((by & 7) << 3) | (bz & 7));
        t.e[h++] = b | (d << 10) | (off << 15);
      }
  return t;
}
static __constant__ AdvectHalo c_advect_halo = make_advect_halo();

// Absent neighbours read the background (leaf 0 is read and discarded, so
// the load does not wait on the branch).  Each thread's table entries, then
// its values, are all loaded before any is used (every thread loads, past
// the end of the table at entry 0), so the round trips overlap instead of
// following one another.
template <int NT>
__device__ __forceinline__ void stage_halo(const GridView& in, const int* nb, float* brick, int tid) {
  constexpr int HQ = (488 + NT - 1) / NT;
  const float bg = in.background;
  uint32_t e[HQ];
#pragma unroll
  for (int q = 0; q < HQ; ++q) {
    const int h = tid + q * NT;
    e[q] = c_advect_halo.e[h < 488 ? h : 0];
  }
  int src[HQ];
  float v[HQ];
#pragma unroll
  for (int q = 0; q < HQ; ++q) {
    src[q] = nb[(e[q] >> 10) & 31u];
    v[q] = in.leaves[src[q] < 0 ? 0 : src[q]].values[e[q] >> 15];
  }
#pragma unroll
  for (int q = 0; q < HQ; ++q)
    if (tid + q * NT < 488) brick[e[q] & 1023u] = src[q] < 0 ? bg : v[q];
}

// The halo as stage_halo, with the neighbour list in a register: lane d of
// every wave holds entry d (d < 27), read with a lane permute, so the list
// needs no LDS and no barrier before the halo loads.
template <int NT>
__device__ __forceinline__ void stage_halo_reg(const GridView& in, int nbv, float* brick, int tid) {
  constexpr int HQ = (488 + NT - 1) / NT;
  const float bg = in.background;
  uint32_t e[HQ];
#pragma unroll
  for (int q = 0; q < HQ; ++q) {
    const int h = tid + q * NT;
    e[q] = c_advect_halo.e[h < 488 ? h : 0];
  }
  int src[HQ];
  float v[HQ];
#pragma unroll
  for (int q = 0; q < HQ; ++q) {
    src[q] = __shfl(nbv, (int)((e[q] >> 10) & 31u));
    v[q] = in.leaves[src[q] < 0 ? 0 : src[q]].values[e[q] >> 15];
  }
#pragma unroll
  for (int q = 0; q < HQ; ++q)
    if (tid + q * NT < 488) brick[e[q] & 1023u] = src[q] < 0 ? bg : v[q];
}

// LDS of one advecting block (5088 bytes without the vortex tables, so 32
// one-wave blocks fit a CU's 160 KB).  TAB: room for the single-vortex row
// tables (sin x, cos x, sin y, cos y, 8 each).
template <bool TAB>
struct AdvectLDST {
  float brick[1000];
  uint16_t list[512];
  uint32_t sgn[16];  // sign mask of the results (SIGNS)
  double vtab[TAB ? 32 : 1];
};
// Whether kernels for velocity kind vk (negative: kinds mixed at run time)
// may run the single-vortex path.
constexpr bool advect_tab(int vk) { return vk == VK_VORTEX || vk < 0; }
template <int VK>
using AdvectLDS = AdvectLDST<advect_tab(VK)>;

// Advects leaf li of `in` into out[li] with NW waves (the whole block,
// tid in [0, 64 NW)): the active voxels are listed by a ballot/popcount
// prefix over the 8 mask words and spread densely over the block's lanes
// (one pass for a band leaf with 4 waves); each result is stored straight
// into the output leaf, whose inactive positions get the background.
// Output origin and mask are the input's (the topology does not change,
// src/tools/advection.rs:34-53).  The background stores follow the
// results, so each output cache line is written within a short window and
// leaves the L2 once.  SIGNS: also writes the leaf's sign mask for `iso`
// to sg[0..15] (bit (o & 31) of word o >> 5 = value at offset o < iso,
// background at the inactive offsets), gathered with LDS atomics as the
// results come, so marching cubes over the output can classify cells
// without reading values.
template <int VK, int SCHEME, int NW, bool SIGNS = false, class LDS>
__device__ __forceinline__ void advect_leaf(const GridView& in, const VelView& vel, NanoLeaf* out, int li,
                                            LDS& s, int tid, double dt, const double tf[3],
                                            uint32_t* sg = nullptr, float iso = 0.0f) {
  constexpr int NT = 64 * NW;
  const int lane = tid & 63, wave = tid >> 6;
  const NanoLeaf& L = in.leaves[li];
  NanoLeaf& O = out[li];
  const float bg = in.background;
  // The independent loads go out first (every lane loads, so no branch
  // waits on them): the neighbour list entry and the leaf's own values.
  constexpr int OQ = (128 + NT - 1) / NT;
  // every wave holds the neighbour list in lanes 0..26 (stage_halo_reg)
  const int nbv = in.nbr[(size_t)li * VR_NBR + (lane < VR_NBR ? lane : 0)];
  const uint64_t mword = L.mask[tid & 7];  // copied to the output by threads 0..7
  const uint32_t sword = SIGNS ? reinterpret_cast<const uint32_t*>(L.mask)[tid & 15] : 0u;
  float4 own[OQ];
#pragma unroll
  for (int q = 0; q < OQ; ++q) {
    const int e = tid + q * NT;
    own[q] = reinterpret_cast<const float4*>(L.values)[e < 128 ? e : 0];
  }
  uint64_t mk[8];  // all eight mask words in one round trip
#pragma unroll
  for (int x = 0; x < 8; ++x) mk[x] = L.mask[x];
  int total = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const uint64_t m = mk[x];
    // active voxels of this slab in lanes below this one (mbcnt: popcount of m under the lane mask)
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if ((x % NW) == wave && ((m >> lane) & 1ull)) s.list[total + rank] = (uint16_t)(x * 64 + lane);
    total += __popcll(m);
  }
  const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
  const bool tab = VK == VK_VORTEX && in.diag;
  if (tab && tid < 16) {
    // row coordinates as view_index_to_world's diagonal form computes them
    const int a = tid >> 3, l = tid & 7;
    const double wc = a == 0 ? in.mat[0] * (double)(ox + l) + in.mat[3] : in.mat[5] * (double)(oy + l) + in.mat[7];
    double sn, cs;
    sincospi_det(wc, &sn, &cs);
    s.vtab[16 * a + l] = sn;
    s.vtab[16 * a + 8 + l] = cs;
  }
#pragma unroll
  for (int q = 0; q < OQ; ++q) {
    const int e = tid + q * NT;
    if (e < 128) {
      const int x = e >> 4, y = (e >> 1) & 7, z = (e & 1) * 4;
      float* b = s.brick + (x + 1) * 100 + (y + 1) * 10 + (z + 1);
      b[0] = own[q].x;
      b[1] = own[q].y;
      b[2] = own[q].z;
      b[3] = own[q].w;
    }
  }
  // Output header: origin and mask are the input's.
  if (tid < 8) O.mask[tid] = mword;
  if (SIGNS && tid < 16) s.sgn[tid] = bg < iso ? ~sword : 0u;
  if (tid < 3) O.origin[tid] = tid == 0 ? ox : (tid == 1 ? oy : oz);
  if (tid == 3) O.pad = 0;
  stage_halo_reg<NT>(in, nbv, s.brick, tid);
  __syncthreads();
  for (int i = tid; i < total; i += NT) {
    const int off = s.list[i];
    const float r = advect_voxel<VK, SCHEME>(in, vel, s.brick, ox, oy, oz, off >> 6, (off >> 3) & 7, off & 7, dt,
                                             tf, tab, s.vtab);
    O.values[off] = r;
    if (SIGNS && r < iso) atomicOr(&s.sgn[off >> 5], 1u << (off & 31));
  }
#pragma unroll
  for (int x = 0; x < 8; ++x)
    if ((x % NW) == wave && !((mk[x] >> lane) & 1ull)) O.values[x * 64 + lane] = bg;
  __syncthreads();
  if (SIGNS && tid < 16) sg[tid] = s.sgn[tid];
}
