// Per-voxel and per-leaf advection device code shared by the single-grid
// kernel (vr_kernels.hip) and the batched persistent kernel (vr_batch.hip).
#pragma once

#include "vr_device.h"

// Trilinear sample of the input at the departure point whose cell is
// (i, j, k) and fractions (fx, fy, fz) (src/tools/interpolation.rs:11-48):
// the cell's corners from the LDS brick of the leaf at (ox, oy, oz) when
// they all lie in it, else one at a time through the tree.
__device__ __forceinline__ float sample_departure(const GridView& in, const float* brick, int ox, int oy, int oz, int i,
                                                  int j, int k, float fx, float fy, float fz) {
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

// Component of the departure point of a uniform velocity u: departure_from's
// arithmetic on one axis (every stage samples the same u), the same
// operations in the same order.
template <int SCHEME>
__device__ __forceinline__ double departure_1d(double x, double u, double dt, const double tf[3]) {
  if constexpr (SCHEME == 1) {
    return x - dt * (u * tf[0]);
  } else {
    const double y1 = x - dt * (u * tf[0]);
    const double y2 = 0.75 * x + 0.25 * (y1 - dt * (u * tf[1]));
    return (1.0 / 3.0) * x + (2.0 / 3.0) * (y2 - dt * (u * tf[2]));
  }
}

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
  return sample_departure(in, brick, ox, oy, oz, f64_to_i32(flx), f64_to_i32(fly), f64_to_i32(flz), fx, fy, fz);
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

// The halo as load tasks derived from the task number, with no table: 72
// four-value tasks (the +-x and +-y faces and the four z-parallel edges,
// four values contiguous along z in the neighbour leaf, one 16-byte load
// each) and 200 one-value tasks (the +-z faces, the x- and y-parallel edges
// and the corners).  A task names its neighbour slot (dx + 1) * 9 + (dy + 1)
// * 3 + (dz + 1), the voxel offset in that neighbour (local coordinate 7 on
// a - side, 0 on a + side) and the brick position.
struct HaloTask {
  int slot, src, dst;
};
__device__ __forceinline__ HaloTask halo_vec_task(int v) {  // v in [0, 72)
  HaloTask t;
  const int h4 = (v & 1) * 4;
  if (v < 64) {
    const int face = v >> 4, r = (v >> 1) & 7;
    if (face < 2) {  // -x (slot 4, its x = 7 slab), +x (slot 22, x = 0): row y = r
      t.slot = face ? 22 : 4;
      t.src = (face ? 0 : 448) + r * 8 + h4;
      t.dst = (face ? 900 : 0) + (r + 1) * 10 + 1 + h4;
    } else {  // -y (slot 10, y = 7), +y (slot 16, y = 0): row x = r
      t.slot = face == 3 ? 16 : 10;
      t.src = r * 64 + (face == 3 ? 0 : 56) + h4;
      t.dst = (r + 1) * 100 + (face == 3 ? 90 : 0) + 1 + h4;
    }
  } else {  // z-parallel edges: x and y at -1 / 8
    const int e = (v - 64) >> 1, px = e >> 1, py = e & 1;
    t.slot = px * 18 + py * 6 + 1;
    t.src = (px ? 0 : 448) + (py ? 0 : 56) + h4;
    t.dst = (px ? 900 : 0) + (py ? 90 : 0) + 1 + h4;
  }
  return t;
}
__device__ __forceinline__ HaloTask halo_one_task(int s) {  // s in [0, 200)
  HaloTask t;
  if (s < 128) {  // -z (slot 12, z = 7), +z (slot 14, z = 0)
    const int f = s >> 6, x = (s >> 3) & 7, y = s & 7;
    t.slot = f ? 14 : 12;
    t.src = x * 64 + y * 8 + (f ? 0 : 7);
    t.dst = (x + 1) * 100 + (y + 1) * 10 + (f ? 9 : 0);
  } else if (s < 192) {
    const int e = ((s - 128) >> 3) & 3, k = s & 7, p1 = e >> 1, p2 = e & 1;
    if (s < 160) {  // y-parallel edges: x and z at -1 / 8
      t.slot = p1 * 18 + 3 + p2 * 2;
      t.src = (p1 ? 0 : 448) + k * 8 + (p2 ? 0 : 7);
      t.dst = (p1 ? 900 : 0) + (k + 1) * 10 + (p2 ? 9 : 0);
    } else {  // x-parallel edges: y and z at -1 / 8
      t.slot = 9 + p1 * 6 + p2 * 2;
      t.src = k * 64 + (p1 ? 0 : 56) + (p2 ? 0 : 7);
      t.dst = (k + 1) * 100 + (p1 ? 90 : 0) + (p2 ? 9 : 0);
    }
  } else {  // corners
    const int c = s - 192, px = (c >> 2) & 1, py = (c >> 1) & 1, pz = c & 1;
    t.slot = px * 18 + py * 6 + pz * 2;
    t.src = (px ? 0 : 448) + (py ? 0 : 56) + (pz ? 0 : 7);
    t.dst = (px ? 900 : 0) + (py ? 90 : 0) + (pz ? 9 : 0);
  }
  return t;
}

// Stages halo tasks with NT threads: NVEC four-value tasks vtask(v) and
// NONE one-value tasks stask(s) (a task with dst < 0 is skipped: its brick
// positions are left unwritten for a caller that never reads them);
// nbget(slot) gives the neighbouring leaf (-1: none).  Every load is issued
// before any LDS write; an absent neighbour reads leaf 0 and is replaced by
// the background.
template <int NT, int NVEC, int NONE, class NB, class VT, class ST>
__device__ __forceinline__ void stage_tasks(const GridView& in, NB nbget, float* brick, int tid, VT vtask,
                                            ST stask) {
  constexpr int VQ = (NVEC + NT - 1) / NT, SQ = (NONE + NT - 1) / NT;
  __builtin_assume(tid >= 0 && tid < NT);
  const float bg = in.background;
  float4 vv[VQ];
  int vsrc[VQ], vdst[VQ];
#pragma unroll
  for (int q = 0; q < VQ; ++q) {
    const int v = tid + q * NT;
    const HaloTask t = vtask(v < NVEC ? v : 0);
    vsrc[q] = nbget(t.slot);
    vdst[q] = v < NVEC ? t.dst : -1;
    vv[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (vdst[q] >= 0)
      vv[q] = *reinterpret_cast<const VR_GLOBAL float4*>(&in.leaves[vsrc[q] < 0 ? 0 : vsrc[q]].values[t.src]);
  }
  float sv[SQ];
  int ssrc[SQ], sdst[SQ];
#pragma unroll
  for (int q = 0; q < SQ; ++q) {
    const int s = tid + q * NT;
    const HaloTask t = stask(s < NONE ? s : 0);
    ssrc[q] = nbget(t.slot);
    sdst[q] = s < NONE ? t.dst : -1;
    sv[q] = 0.0f;
    if (sdst[q] >= 0) sv[q] = in.leaves[ssrc[q] < 0 ? 0 : ssrc[q]].values[t.src];
  }
#pragma unroll
  for (int q = 0; q < VQ; ++q)
    if (vdst[q] >= 0) {
      const float4 a = vsrc[q] < 0 ? make_float4(bg, bg, bg, bg) : vv[q];
      float* b = brick + vdst[q];
      b[0] = a.x;
      b[1] = a.y;
      b[2] = a.z;
      b[3] = a.w;
    }
#pragma unroll
  for (int q = 0; q < SQ; ++q)
    if (sdst[q] >= 0) brick[sdst[q]] = ssrc[q] < 0 ? bg : sv[q];
}

// The 488 halo values of the 10^3 brick (the positions with a local
// coordinate -1 or 8), of the neighbour slots in `slots` (bit per slot).
template <int NT, class NB>
__device__ __forceinline__ void stage_halo_tasks(const GridView& in, NB nbget, float* brick, int tid,
                                                 uint32_t slots = 0x7FFFFFFu) {
  stage_tasks<NT, 72, 200>(
      in, nbget, brick, tid,
      [slots](int v) {
        HaloTask t = halo_vec_task(v);
        if (!((slots >> t.slot) & 1u)) t.dst = -1;
        return t;
      },
      [slots](int s) {
        HaloTask t = halo_one_task(s);
        if (!((slots >> t.slot) & 1u)) t.dst = -1;
        return t;
      });
}

// The halo on one side per axis, sx, sy, sz in {-1, 0, +1} (0: that axis
// needs none): the face of each chosen side, the edges and the corner where
// two or three chosen sides meet -- at most 217 values, as 34 four-value
// tasks (the x and y faces by rows along z, the z-parallel edge) and 81
// one-value tasks (the z face, the y- and x-parallel edges, the corner).
// Same slots, sources and brick positions as the full task set's.
template <int NT, class NB>
__device__ __forceinline__ void stage_halo_sided(const GridView& in, NB nbget, float* brick, int tid, int sx, int sy,
                                                 int sz) {
  // per side: source offset of the neighbour's facing slab and brick base
  const int xs = sx < 0 ? 448 : 0, xd = sx < 0 ? 0 : 900, ys = sy < 0 ? 56 : 0, yd = sy < 0 ? 0 : 90;
  const int zs = sz < 0 ? 7 : 0, zd = sz < 0 ? 0 : 9;
  stage_tasks<NT, 34, 81>(
      in, nbget, brick, tid,
      [=](int v) {
        const int h4 = (v & 1) * 4, r = (v >> 1) & 7;
        HaloTask t{0, 0, -1};
        if (v < 16) {  // x face, row y = r
          t = HaloTask{(sx + 1) * 9 + 4, xs + r * 8 + h4, xd + (r + 1) * 10 + 1 + h4};
          if (sx == 0) t.dst = -1;
        } else if (v < 32) {  // y face, row x = r
          t = HaloTask{9 + (sy + 1) * 3 + 1, r * 64 + ys + h4, (r + 1) * 100 + yd + 1 + h4};
          if (sy == 0) t.dst = -1;
        } else {  // z-parallel edge
          t = HaloTask{(sx + 1) * 9 + (sy + 1) * 3 + 1, xs + ys + h4, xd + yd + 1 + h4};
          if (sx == 0 || sy == 0) t.dst = -1;
        }
        return t;
      },
      [=](int s) {
        HaloTask t{0, 0, -1};
        if (s < 64) {  // z face, row (x, y) = (s >> 3, s & 7)
          t = HaloTask{12 + sz + 1, (s >> 3) * 64 + (s & 7) * 8 + zs, ((s >> 3) + 1) * 100 + ((s & 7) + 1) * 10 + zd};
          if (sz == 0) t.dst = -1;
        } else if (s < 72) {  // y-parallel edge, y = s - 64
          t = HaloTask{(sx + 1) * 9 + 3 + sz + 1, xs + (s - 64) * 8 + zs, xd + (s - 63) * 10 + zd};
          if (sx == 0 || sz == 0) t.dst = -1;
        } else if (s < 80) {  // x-parallel edge, x = s - 72
          t = HaloTask{9 + (sy + 1) * 3 + sz + 1, (s - 72) * 64 + ys + zs, (s - 71) * 100 + yd + zd};
          if (sy == 0 || sz == 0) t.dst = -1;
        } else {  // corner
          t = HaloTask{(sx + 1) * 9 + (sy + 1) * 3 + sz + 1, xs + ys + zs, xd + yd + zd};
          if (sx == 0 || sy == 0 || sz == 0) t.dst = -1;
        }
        return t;
      });
}

// The halo with the neighbour list in LDS (nb[d], d < 27), after a barrier.
template <int NT>
__device__ __forceinline__ void stage_halo(const GridView& in, const int* nb, float* brick, int tid) {
  stage_halo_tasks<NT>(in, [nb](int d) { return nb[d]; }, brick, tid);
}

// The halo with the neighbour list in a register: lane d of every wave
// holds entry d (d < 27), read with a lane permute, so the list needs no
// LDS and no barrier before the halo loads.
template <int NT>
__device__ __forceinline__ void stage_halo_reg(const GridView& in, int nbv, float* brick, int tid,
                                               uint32_t slots = 0x7FFFFFFu) {
  stage_halo_tasks<NT>(in, [nbv](int d) { return __shfl(nbv, d); }, brick, tid, slots);
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

// The loads that start leaf li, all vector loads with no dependence among
// them: its 512 values (own), neighbour list entry lane (nbv, lanes < 27)
// and record header dword lane (hdr, lanes < 20: origin, pad and the 8
// mask words), so that a caller can issue them a leaf ahead.
template <int NW>
struct AdvectLoads {
  float4 own[(128 + 64 * NW - 1) / (64 * NW)];
  int nbv;
  uint32_t hdr;
};
template <int NW>
__device__ __forceinline__ AdvectLoads<NW> advect_issue(const VR_GLOBAL NanoLeaf* leaves, const VR_GLOBAL int* nbr,
                                                        int li, int tid) {
  constexpr int NT = 64 * NW;
  const int lane = tid & 63;
  const VR_GLOBAL NanoLeaf* L = leaves + li;
  AdvectLoads<NW> a;
  a.hdr = reinterpret_cast<const VR_GLOBAL uint32_t*>(L)[lane < 20 ? lane : 0];
  a.nbv = nbr[(size_t)li * VR_NBR + (lane < VR_NBR ? lane : 0)];
#pragma unroll
  for (int q = 0; q < (128 + NT - 1) / NT; ++q) {
    const int e = tid + q * NT;
    a.own[q] = reinterpret_cast<const VR_GLOBAL float4*>(L->values)[e < 128 ? e : 0];
  }
  return a;
}

// Advects leaf li of `in` into out[li] with NW waves (the whole block,
// tid in [0, 64 NW)) from its loads `ld` (advect_issue): the active voxels are listed by a ballot/popcount
// prefix over the 8 mask words and spread densely over the block's lanes
// (one pass for a band leaf with 4 waves); the results are gathered in the
// brick's interior and leave, with the background at the inactive
// positions, as the whole output leaf in 16-byte stores.
// Output origin and mask are the input's (the topology does not change,
// src/tools/advection.rs:34-53).  SIGNS: also writes the leaf's sign mask for `iso`
// to sg[0..15] (bit (o & 31) of word o >> 5 = value at offset o < iso,
// background at the inactive offsets), gathered with LDS atomics as the
// results come, so marching cubes over the output can classify cells
// without reading values.
template <int VK, int SCHEME, int NW, bool SIGNS = false, class LDS>
__device__ __forceinline__ void advect_leaf_loaded(const GridView& in, const VelView& vel, NanoLeaf* out, int li,
                                                   const AdvectLoads<NW>& ld, LDS& s, int tid, double dt,
                                                   const double tf[3], uint32_t* sg = nullptr, float iso = 0.0f) {
  constexpr int NT = 64 * NW;
  const int lane = tid & 63, wave = tid >> 6;
  NanoLeaf& O = out[li];
  const float bg = in.background;
  constexpr int OQ = (128 + NT - 1) / NT;
  const int nbv = ld.nbv;  // every wave holds the neighbour list in lanes 0..26 (stage_halo_reg)
  uint64_t mk[8];
#pragma unroll
  for (int x = 0; x < 8; ++x)
    mk[x] = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ld.hdr, 4 + 2 * x) |
            ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ld.hdr, 5 + 2 * x) << 32);
  int total = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const uint64_t m = mk[x];
    // active voxels of this slab in lanes below this one (mbcnt: popcount of m under the lane mask)
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if ((x % NW) == wave && ((m >> lane) & 1ull)) s.list[total + rank] = (uint16_t)(x * 64 + lane);
    total += __popcll(m);
  }
  const int ox = __builtin_amdgcn_readlane((int)ld.hdr, 0), oy = __builtin_amdgcn_readlane((int)ld.hdr, 1),
            oz = __builtin_amdgcn_readlane((int)ld.hdr, 2);
  const bool tab = VK == VK_VORTEX && in.diag;
  // Uniform velocity on a diagonal transform: component a of a departure
  // point depends on the voxel's coordinate on axis a alone, so the cell
  // index and fraction on each axis are computed once per leaf for the 8
  // local coordinates, lane 8 a + l holding those of coordinate l on axis a
  // (lanes from 24 on repeat axis 2), and each voxel gathers its three with
  // lane permutes.
  // The tables also tell which sides of the brick the departure cells
  // reach: the halo on the - side of axis a is read only when some cell
  // starts at local -1, the + side only when one starts at 7 (a cell
  // outside [-1, 7] reads through the tree), so only the neighbour slots
  // (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1) whose sides are reached (bits
  // of `slots`) are loaded; under a CFL number below one that is at most
  // one side per axis.
  // When each axis reaches at most one side, only those sides are staged
  // (stage_halo_sided: about half the tasks).
  const bool sep = VK == VK_UNIFORM && in.diag;
  int sep_i = 0;
  float sep_f = 0.0f;
  uint32_t slots = 0x7FFFFFFu;
  bool sided = false;
  int sd[3] = {0, 0, 0};
  if (sep) {
    const int l = lane & 7, ax = lane < 8 ? 0 : (lane < 16 ? 1 : 2);
    // the lane's axis parameters, selected among values that an empty asm
    // makes opaque: a select among loads of the views would become a load at
    // a lane-dependent subscript, which puts the views in scratch memory
    double p[9] = {in.mat[0], in.mat[5], in.mat[10], in.mat[3], in.mat[7], in.mat[11], vel.a[0], vel.a[1], vel.a[2]};
#pragma unroll
    for (int k = 0; k < 9; ++k) asm volatile("" : "+v"(p[k]));
    const double m = ax == 0 ? p[0] : (ax == 1 ? p[1] : p[2]);
    const double c = ax == 0 ? p[3] : (ax == 1 ? p[4] : p[5]);
    const double u = ax == 0 ? p[6] : (ax == 1 ? p[7] : p[8]);
    const int o = ax == 0 ? ox : (ax == 1 ? oy : oz);
    const double ix = departure_1d<SCHEME>(m * (double)(o + l) + c, u, dt, tf) * in.inv_vs;
    const double fl = floor(ix);
    sep_f = (float)(ix - fl);
    sep_i = f64_to_i32(fl);
    const int b = sep_i - o;
    const uint64_t lo = __ballot(b == -1), hi = __ballot(b == 7);
    uint32_t side[3];  // bit d + 1: side d (-1, 0, +1) of the axis is reached
#pragma unroll
    for (int a = 0; a < 3; ++a)
      side[a] = (((lo >> (8 * a)) & 0xFFu) ? 1u : 0u) | 2u | (((hi >> (8 * a)) & 0xFFu) ? 4u : 0u);
    const uint32_t yz = ((side[1] & 1u) ? side[2] : 0u) | (side[2] << 3) | ((side[1] & 4u) ? side[2] << 6 : 0u);
    slots = ((side[0] & 1u) ? yz : 0u) | (yz << 9) | ((side[0] & 4u) ? yz << 18 : 0u);
    sided = side[0] != 7u && side[1] != 7u && side[2] != 7u;
#pragma unroll
    for (int a = 0; a < 3; ++a) sd[a] = (side[a] & 1u) ? -1 : ((side[a] & 4u) ? 1 : 0);
  }
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
      b[0] = ld.own[q].x;
      b[1] = ld.own[q].y;
      b[2] = ld.own[q].z;
      b[3] = ld.own[q].w;
    }
  }
  // Output header: origin and mask are the input's, the pad word 0.
  if (tid < 20) reinterpret_cast<uint32_t*>(&O)[tid] = tid == 3 ? 0u : ld.hdr;
  if (SIGNS) {
    const uint32_t sword = (uint32_t)__shfl((int)ld.hdr, 4 + (lane & 15));  // mask dword lane & 15
    if (tid < 16) s.sgn[tid] = bg < iso ? ~sword : 0u;
  }
  if (sided)
    stage_halo_sided<NT>(in, [nbv](int d) { return __shfl(nbv, d); }, s.brick, tid, sd[0], sd[1], sd[2]);
  else
    stage_halo_reg<NT>(in, nbv, s.brick, tid, slots);
  __syncthreads();
  // Results stay in registers until every voxel has sampled the brick
  // (at most 512 / NT per thread), then they go into the brick's interior,
  // and the output leaf leaves with 16-byte stores: two per lane of a
  // one-wave block instead of one store per list pass for the results and
  // one per slab for the background.
  constexpr int RQ = 512 / NT;
  float r[RQ];
  const int wfirst = __builtin_amdgcn_readfirstlane(wave * 64);
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    r[q] = 0.0f;
    if (wfirst + q * NT >= total) continue;  // no voxel of this wave in this pass (wave-uniform)
    const int i = tid + q * NT;
    const int off = s.list[i < total ? i : 0];
    const int vx = off >> 6, vy = (off >> 3) & 7, vz = off & 7;
    if (sep) {
      // the whole wave permutes (a lane past the list reads harmlessly)
      const int ci = __shfl(sep_i, vx), cj = __shfl(sep_i, 8 + vy), ck = __shfl(sep_i, 16 + vz);
      const float fx = __shfl(sep_f, vx), fy = __shfl(sep_f, 8 + vy), fz = __shfl(sep_f, 16 + vz);
      if (i < total) r[q] = sample_departure(in, s.brick, ox, oy, oz, ci, cj, ck, fx, fy, fz);
    } else if (i < total) {
      r[q] = advect_voxel<VK, SCHEME>(in, vel, s.brick, ox, oy, oz, vx, vy, vz, dt, tf, tab, s.vtab);
    }
    if (SIGNS && i < total && r[q] < iso) atomicOr(&s.sgn[off >> 5], 1u << (off & 31));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int i = tid + q * NT;
    if (i < total) {
      const int off = s.list[i];
      s.brick[((off >> 6) + 1) * 100 + (((off >> 3) & 7) + 1) * 10 + (off & 7) + 1] = r[q];
    }
  }
  __syncthreads();
  // The output leaf: the results where the mask is set, the background
  // elsewhere (the brick still holds the input there).  A float4's four
  // mask bits are byte y & 3 of header dword 4 + 2 x + y / 4, fetched by
  // the whole wave with a lane permute.
#pragma unroll
  for (int q = 0; q < OQ; ++q) {
    const int e = tid + q * NT;
    const int x = (e >> 4) & 7, y = (e >> 1) & 7, z = (e & 1) * 4;
    const uint32_t w = (uint32_t)__shfl((int)ld.hdr, 4 + 2 * x + (y >> 2));
    if (e < 128) {
      const uint32_t act = w >> ((y & 3) * 8 + z);
      const float* b = s.brick + (x + 1) * 100 + (y + 1) * 10 + (z + 1);
      reinterpret_cast<float4*>(O.values)[e] = make_float4((act & 1u) ? b[0] : bg, (act & 2u) ? b[1] : bg,
                                                           (act & 4u) ? b[2] : bg, (act & 8u) ? b[3] : bg);
    }
  }
  if (SIGNS && tid < 16) sg[tid] = s.sgn[tid];
  // the next leaf of a block that loops over several restages the brick
  __syncthreads();
}

// advect_leaf_loaded with the leaf's loads issued here.
template <int VK, int SCHEME, int NW, bool SIGNS = false, class LDS>
__device__ __forceinline__ void advect_leaf(const GridView& in, const VelView& vel, NanoLeaf* out, int li,
                                            LDS& s, int tid, double dt, const double tf[3],
                                            uint32_t* sg = nullptr, float iso = 0.0f) {
  advect_leaf_loaded<VK, SCHEME, NW, SIGNS>(in, vel, out, li, advect_issue<NW>(in.leaves, in.nbr, li, tid), s, tid,
                                            dt, tf, sg, iso);
}
