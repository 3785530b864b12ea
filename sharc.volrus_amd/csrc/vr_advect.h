// Per-voxel and per-leaf advection device code shared by the single-grid
// kernel (vr_kernels.hip) and the batched persistent kernel (vr_batch.hip).
#pragma once

#include "vr_device.h"

// New value of the active voxel (ox + x, oy + y, oz + z): back-trace, then
// trilinear sampling of the input (src/tools/advection.rs:39-53).
// Departure-cell corners inside the LDS brick (leaf plus one-voxel halo,
// 10^3) come from LDS; any other corner goes through the root/internal
// lookup, so any dt is handled.
template <int VK, int SCHEME>
__device__ __forceinline__ float advect_voxel(const GridView& in, const VelView& vel, const float* brick, int ox,
                                              int oy, int oz, int x, int y, int z, double dt, const double tf[3]) {
  double w[3], p[3];
  index_to_world(in.mat, ox + x, oy + y, oz + z, w);
  departure<VK, SCHEME>(vel, w, dt, tf, p);
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
#pragma unroll
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

// Halo: the 1000 - 512 brick positions with some local coordinate -1 or 8.
template <int NT>
__device__ __forceinline__ void stage_halo(const GridView& in, const int* nb, float* brick, int tid) {
  for (int h = tid; h < 488; h += NT) {
    int bx, by, bz;
    if (h < 200) {  // x = -1 or x = 8 planes (10 x 10 each)
      bx = h < 100 ? -1 : 8;
      const int r = h % 100;
      by = r / 10 - 1;
      bz = r % 10 - 1;
    } else if (h < 360) {  // y = -1 or y = 8 rows, x in 0..7 (8 x 10 each)
      const int r = h - 200;
      by = r < 80 ? -1 : 8;
      const int u = r % 80;
      bx = u / 10;
      bz = u % 10 - 1;
    } else {  // z = -1 or z = 8, x, y in 0..7 (8 x 8 each)
      const int r = h - 360;
      bz = r < 64 ? -1 : 8;
      const int u = r & 63;
      bx = u >> 3;
      by = u & 7;
    }
    const int dx = bx < 0 ? -1 : (bx > 7 ? 1 : 0), dy = by < 0 ? -1 : (by > 7 ? 1 : 0),
              dz = bz < 0 ? -1 : (bz > 7 ? 1 : 0);
    const int src = nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
    brick[(bx + 1) * 100 + (by + 1) * 10 + (bz + 1)] =
        src < 0 ? in.background : in.leaves[src].values[voxel_offset(bx, by, bz)];
  }
}

// LDS of one advecting block.
struct AdvectLDS {
  float brick[1000];
  float res[512];
  uint16_t list[512];
  int nb[VR_NBR];
};

// Advects leaf li of `in` into out[li] with NW waves (the whole block,
// tid in [0, 64 NW)): the active voxels are listed by a ballot/popcount
// prefix over the 8 mask words and spread densely over the block's lanes
// (one pass for a band leaf with 4 waves), then the leaf is written back
// slab by slab, coalesced, with the background in the inactive positions.
// Output origin and mask are the input's (the topology does not change,
// src/tools/advection.rs:34-53).
template <int VK, int SCHEME, int NW>
__device__ __forceinline__ void advect_leaf(const GridView& in, const VelView& vel, NanoLeaf* out, int li,
                                            AdvectLDS& s, int tid, double dt, const double tf[3]) {
  constexpr int NT = 64 * NW;
  const int lane = tid & 63, wave = tid >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const NanoLeaf& L = in.leaves[li];
  if (tid < VR_NBR) s.nb[tid] = in.nbr[(size_t)li * VR_NBR + tid];
  uint64_t m[8];
  int pre[8];
  int total = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    m[x] = L.mask[x];
    pre[x] = total;
    total += __popcll(m[x]);
  }
#pragma unroll
  for (int x = 0; x < 8; ++x)
    if ((x % NW) == wave && ((m[x] >> lane) & 1ull)) s.list[pre[x] + __popcll(m[x] & below)] = (uint16_t)(x * 64 + lane);
  stage_interior<NT>(in, li, s.brick, tid);
  __syncthreads();
  stage_halo<NT>(in, s.nb, s.brick, tid);
  __syncthreads();
  const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
  for (int i = tid; i < total; i += NT) {
    const int off = s.list[i];
    s.res[off] = advect_voxel<VK, SCHEME>(in, vel, s.brick, ox, oy, oz, off >> 6, (off >> 3) & 7, off & 7, dt, tf);
  }
  __syncthreads();
  NanoLeaf& O = out[li];
  if (tid < 3) O.origin[tid] = L.origin[tid];
  if (tid == 3) O.pad = 0;
  if (tid < 8) O.mask[tid] = m[tid];
  const float bg = in.background;
#pragma unroll
  for (int x = 0; x < 8; ++x)
    if ((x % NW) == wave) {
      // Both operands are read, so the select is of values, not of an LDS
      // and a global address (which would need a FLAT load).
      const float r = s.res[x * 64 + lane];
      O.values[x * 64 + lane] = ((m[x] >> lane) & 1ull) ? r : bg;
    }
  __syncthreads();
}
