// gfx950 kernels for the narrow-band level-set path.
//
// One wave64 per 8^3 leaf (or per marching-cubes tile): lane = y * 8 + z and
// each lane walks the 8 x-slabs, so the active mask word for slab x is the
// wave's ballot layout (mask bit = x * 64 + lane, src/tree/leaf.rs:52-56).
// Leaf bricks plus halo are staged in LDS; every count and offset is made by
// wave/block prefix scans, so outputs are deterministic.
#include <cstdlib>

#include "vr_advect.h"
#include "vr_kernels.h"

// ------------------------------------------------------------------ small utilities
__global__ void k_fill_int(int* p, int v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// child[r * 4096 + slot] = leaf index, origins[i] = leaf origin.
__global__ void k_build_child(const NanoLeaf* leaves, const int* count, const int4* roots, int root_count, int* child,
                              int4* origins) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int x = leaves[i].origin[0], y = leaves[i].origin[1], z = leaves[i].origin[2];
    origins[i] = make_int4(x, y, z, 0);
    const int kx = x & ~127, ky = y & ~127, kz = z & ~127;
    int lo = 0, hi = root_count;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int4 r = roots[mid];
      if (coord_less(r.x, r.y, r.z, kx, ky, kz)) lo = mid + 1; else hi = mid;
    }
    const int slot = (((x - kx) >> 3) << 8) | (((y - ky) >> 3) << 4) | ((z - kz) >> 3);
    child[(size_t)lo * 4096 + slot] = i;
  }
}

__global__ void k_build_nbr(GridView g, int* nbr) {
  const int n = *g.count;
  const size_t total = (size_t)n * VR_NBR;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / VR_NBR), d = (int)(t % VR_NBR);
    const int4 o = g.origins[i];
    const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
    nbr[t] = find_leaf(g, o.x + 8 * dx, o.y + 8 * dy, o.z + 8 * dz);
  }
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n ints in three launches: per-block scan of 1024
// elements, a single-block scan of the block sums, then the add-back.
__device__ __forceinline__ int block_excl_scan_256(int v, int* s_wave, int* total) {
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int incl = wave_incl_scan(v);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int s = s_wave[w];
    if (w < wave) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Small inputs (n <= VR_SCAN_SINGLE_MAX) in one launch: one block of 1024
// threads walks the input in tiles of 4096 (4 consecutive elements per
// thread) with a running carry; NA = 2 scans a second array of the same
// length in the same pass.  n_dev (optional): the live length is
// min(n, *n_dev * n_mult), a device-side count below the host bound n.
#define VR_SCAN_SINGLE_MAX 32768
template <int NA>
__global__ __launch_bounds__(1024) void k_scan_single(const int* in0, int* out0, int* total0, const int* in1,
                                                      int* out1, int* total1, int n, const int* n_dev, int n_mult) {
  __shared__ int s_wave[NA][16];
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  if (n_dev) {
    const long long live = (long long)*n_dev * n_mult;
    if (live < n) n = (int)live;
  }
  const int* in[2] = {in0, in1};
  int* out[2] = {out0, out1};
  int carry[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) carry[a] = 0;
  for (int base = 0; base < n; base += 4096) {
    const int i0 = base + (int)threadIdx.x * 4;
    int v[NA][4], sum[NA], incl[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
      for (int k = 0; k < 4; ++k) v[a][k] = i0 + k < n ? in[a][i0 + k] : 0;
      sum[a] = (v[a][0] + v[a][1]) + (v[a][2] + v[a][3]);
      incl[a] = wave_incl_scan(sum[a]);
      if (lane == 63) s_wave[a][wave] = incl[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      int off = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const int ws = s_wave[a][w];
        if (w < wave) off += ws;
        tot += ws;
      }
      int run = carry[a] + off + incl[a] - sum[a];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (i0 + k < n) out[a][i0 + k] = run;
        run += v[a][k];
      }
      carry[a] += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (total0) *total0 = carry[0];
    if (NA == 2 && total1) *total1 = carry[NA - 1];
  }
}

__global__ __launch_bounds__(256) void k_scan_blocks(const int* in, int* out, int* block_sums, size_t n) {
  __shared__ int s_wave[4];
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    s += v[q];
  }
  int total;
  int off = block_excl_scan_256(s, s_wave, &total);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (base + q < n) out[base + q] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_sums(int* sums, int nb, int* total_out) {
  __shared__ int s_wave[4];
  int carry = 0;
  for (int base = 0; base < nb; base += 256) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? sums[i] : 0;
    int tot;
    const int off = block_excl_scan_256(v, s_wave, &tot);
    if (i < nb) sums[i] = carry + off;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ void k_scan_add(int* out, const int* sums, size_t n) {
  const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
  const int add = sums[blockIdx.x];
  for (int q = 0; q < 4; ++q)
    if (i + q * 256 < n) out[i + q * 256] += add;
}

// ------------------------------------------------------------------ advection
// One leaf per block of NW waves (no loop over leaves, so nothing per-leaf
// is kept live across leaves): the active voxels are listed in LDS by a
// ballot/popcount prefix over the 8 mask words and spread over the block's
// lanes; results go straight to the output leaf, the background to its
// inactive positions.  NW = 1 issues the fewest
// instructions per leaf (dense passes of 64 active voxels); more waves
// finish a leaf sooner, for grids with too few leaves to fill the device:
// eight for the analytic fields (every thread stages at most one halo value
// and advects at most one voxel of a band leaf; on the 1/512 single-vortex
// rollout 8 waves beat 2, 4 and 16), four for the sampled fields (dense or
// grids, whose per-voxel work is loads rather than arithmetic; 8 was slower
// on the 1/256 dense-field rollout).
template <int VK, int SCHEME, int NW>
__global__ __launch_bounds__(64 * NW) void k_advect(GridView in, VelView vel, NanoLeaf* out, double dt, double tf0,
                                                    double tf1, double tf2) {
  __shared__ AdvectLDS<VK> s;
  const int n = *in.count;
  const double tf[3] = {tf0, tf1, tf2};
  int li_end, li_step;
  // One leaf per block: the grid has at least ceil(n / 8) blocks per XCD.
  const int li = xcd_first(n, &li_end, &li_step);
  if (li < li_end) advect_leaf<VK, SCHEME, NW>(in, vel, out, li, s, threadIdx.x, dt, tf);
}

template <int VK, int SCHEME>
static hipError_t launch_advect_t(const GridView& in, const VelView& vel, NanoLeaf* out, int cap, double dt,
                                  const double tf[3], int nw, hipStream_t s) {
  const int blocks = 8 * ((cap + 7) / 8 + 1);  // >= ceil(count / 8) blocks on each XCD
  constexpr int NWS = (VK == VK_DENSE || VK == VK_GRIDS3) ? 4 : 8;
  if (nw == 1)
    hipLaunchKernelGGL((k_advect<VK, SCHEME, 1>), dim3(blocks), dim3(64), 0, s, in, vel, out, dt, tf[0], tf[1], tf[2]);
  else
    hipLaunchKernelGGL((k_advect<VK, SCHEME, NWS>), dim3(blocks), dim3(64 * NWS), 0, s, in, vel, out, dt, tf[0], tf[1],
                       tf[2]);
  return hipGetLastError();
}

hipError_t vr_launch_advect(const GridView& in, const VelView& vel, NanoLeaf* out, int cap, double dt,
                            const double tf[3], int scheme, int nw, hipStream_t s) {
#define VR_ADV(VK)                                                                         \
  case VK:                                                                                 \
    return scheme == 3 ? launch_advect_t<VK, 3>(in, vel, out, cap, dt, tf, nw, s)           \
                       : launch_advect_t<VK, 1>(in, vel, out, cap, dt, tf, nw, s);
  switch (vel.kind) {
    VR_ADV(VK_GRIDS3)
    VR_ADV(VK_DENSE)
    VR_ADV(VK_UNIFORM)
    VR_ADV(VK_RIGID)
    VR_ADV(VK_VORTEX)
    default: return hipErrorInvalidValue;
  }
#undef VR_ADV
}

// ------------------------------------------------------------------ zero crossings
// src/tools/level_set_track.rs:14-47: active voxel with value == 0, or with
// an active 6-neighbour whose f32 product with it is < 0.  The leaf and the
// six face layers of its neighbours are staged in LDS with inactive or
// absent voxels as +0: a product with +0 is never < 0, which is the same
// test as skipping them.
// *any is raised to gen when the grid has a crossing.  One leaf per block of
// four waves (wave w takes x = w and w + 4).
__global__ __launch_bounds__(256) void k_crossings(GridView g, uint64_t* cross, int* any, int gen) {
  __shared__ int nb[VR_NBR];
  __shared__ float brick[1000];  // (x + 1) * 100 + (y + 1) * 10 + (z + 1), x, y, z in -1..8
  __shared__ int s_found;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = *g.count;
  const int y = lane >> 3, z = lane & 7;
  int li_end, li_step;
  for (int li = xcd_first(n, &li_end, &li_step); li < li_end; li += li_step) {
    const NanoLeaf& L = g.leaves[li];
    if (tid < VR_NBR) nb[tid] = g.nbr[(size_t)li * VR_NBR + tid];
    if (tid == 0) s_found = 0;
    if (tid < 128) {
      const float4 v = reinterpret_cast<const float4*>(L.values)[tid];
      const int x = tid >> 4, yy = (tid >> 1) & 7, zz = (tid & 1) * 4;
      const uint32_t m = (uint32_t)(L.mask[x] >> (yy * 8 + zz)) & 0xFu;
      float* b = brick + (x + 1) * 100 + (yy + 1) * 10 + (zz + 1);
      b[0] = (m & 1u) ? v.x : 0.0f;
      b[1] = (m & 2u) ? v.y : 0.0f;
      b[2] = (m & 4u) ? v.z : 0.0f;
      b[3] = (m & 8u) ? v.w : 0.0f;
    }
    __syncthreads();
    // face layers: f = 0..5 for x = -1, x = 8, y = -1, y = 8, z = -1, z = 8;
    // a thread's two entries load together (an absent neighbour reads leaf
    // 0, which is discarded)
    float fv[2];
    uint64_t fm[2];
    int fb[2], fsrc[2], foff[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int h = tid + 256 * q < 384 ? tid + 256 * q : tid;
      const int f = h >> 6, u = (h >> 3) & 7, w = h & 7;
      const int side = (f & 1) ? 8 : -1;
      const int bx = f < 2 ? side : u, by = f < 2 ? u : (f < 4 ? side : w), bz = f < 4 ? w : side;
      const int dx = bx < 0 ? -1 : (bx > 7 ? 1 : 0), dy = by < 0 ? -1 : (by > 7 ? 1 : 0),
                dz = bz < 0 ? -1 : (bz > 7 ? 1 : 0);
      fsrc[q] = nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
      foff[q] = voxel_offset(bx, by, bz);
      fb[q] = (bx + 1) * 100 + (by + 1) * 10 + (bz + 1);
      const NanoLeaf& S = g.leaves[fsrc[q] < 0 ? 0 : fsrc[q]];
      fv[q] = S.values[foff[q]];
      fm[q] = S.mask[foff[q] >> 6];
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (tid + 256 * q < 384)
        brick[fb[q]] = (fsrc[q] >= 0 && ((fm[q] >> (foff[q] & 63)) & 1ull)) ? fv[q] : 0.0f;
    __syncthreads();
    bool found = false;
    for (int x = wave; x < 8; x += 4) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      const int c = (x + 1) * 100 + (y + 1) * 10 + (z + 1);
      const float v = brick[c];
      const bool cr = act && (v == 0.0f || v * brick[c + 100] < 0.0f || v * brick[c - 100] < 0.0f ||
                              v * brick[c + 10] < 0.0f || v * brick[c - 10] < 0.0f || v * brick[c + 1] < 0.0f ||
                              v * brick[c - 1] < 0.0f);
      const uint64_t word = __ballot(cr);
      if (lane == 0) cross[(size_t)li * 8 + x] = word;
      found |= word != 0;
    }
    if (found && lane == 0) s_found = 1;
    __syncthreads();
    if (tid == 0 && s_found) atomicMax(any, gen);
  }
}

// ------------------------------------------------------------------ band rebuild
// Bounded-radius form of src/tools/level_set_track.rs:80-114: the crossing
// masks of the 27 neighbouring leaves are staged as 24 x 24 rows of 24 z-bits;
// each active voxel takes the smallest squared lattice distance to a set bit
// within |d| <= R = ceil(half_width) <= 8, which gives the brute-force result
// exactly (kept set and values).  The window is a box and d^2 a sum of
// squares, so the minimum separates exactly: per row the nearest bit along z
// (bit scans), then the minimum over dy, then over dx, in small integers.
// No crossings anywhere: the grid is copied unchanged (the reference returns
// early, :85-87).
// keep[*count .. keep_len) is cleared (block 0) for a scan of keep_len flags.
// active_out (optional) gets += the active voxels of the output.  One leaf
// per block of four waves (wave w takes x = w and w + 4 in the last pass).
__global__ __launch_bounds__(256) void k_rebuild(GridView g, const uint64_t* cross, const int* any, int gen,
                                                NanoLeaf* out, int* keep, int keep_len, int R, float vs, float band,
                                                unsigned long long* active_out) {
  __shared__ uint32_t rows[24 * 24];
  __shared__ uint8_t dz2[24 * 24 * 8];  // row (X, Y), z: dz^2 to the nearest bit, 255 = none within R
  __shared__ uint8_t dyz2[24 * 8 * 8];  // X, y, z: min over dy of dy^2 + dz2, 255 = none
  __shared__ int nb[VR_NBR];
  __shared__ int s_left[4];
  __shared__ unsigned long long s_cnt[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = *g.count;
  if (blockIdx.x == 0)
    for (int i = n + tid; i < keep_len; i += 256) keep[i] = 0;
  const bool have = *any == gen;
  const int y = lane >> 3, z = lane & 7;
  const int lo = 8 - R, span = 8 + 2 * R;  // rows X, Y in [lo, lo + span)
  int li_end, li_step;
  for (int li = xcd_first(n, &li_end, &li_step); li < li_end; li += li_step) {
    const NanoLeaf& L = g.leaves[li];
    NanoLeaf& O = out[li];
    if (tid < 3) O.origin[tid] = L.origin[tid];
    if (tid == 3) O.pad = 0;
    if (!have) {
      if (tid < 8) O.mask[tid] = L.mask[tid];
      for (int i = tid; i < 512; i += 256) O.values[i] = L.values[i];
      if (tid == 0) {
        keep[li] = 1;
        if (active_out) {
          unsigned long long c = 0;
          for (int w = 0; w < 8; ++w) c += (unsigned long long)__popcll(L.mask[w]);
          atomicAdd(active_out, c);
        }
      }
      continue;
    }
    if (tid < VR_NBR) nb[tid] = g.nbr[(size_t)li * VR_NBR + tid];
    __syncthreads();
    {
      // a thread's (up to) three rows of three words all load together (an
      // absent neighbour reads leaf 0's word, which is discarded)
      uint64_t cw[3][3];
      int csrc[3][3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int r = tid + 256 * q < 576 ? tid + 256 * q : tid;
        const int X = r / 24, Y = r % 24, nx = X >> 3, ny = Y >> 3;
#pragma unroll
        for (int nz = 0; nz < 3; ++nz) {
          csrc[q][nz] = nb[nx * 9 + ny * 3 + nz];
          cw[q][nz] = cross[(size_t)(csrc[q][nz] < 0 ? 0 : csrc[q][nz]) * 8 + (X & 7)];
        }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int r = tid + 256 * q;
        if (r < 576) {
          const int Y = r % 24;
          uint32_t bits = 0;
#pragma unroll
          for (int nz = 0; nz < 3; ++nz)
            if (csrc[q][nz] >= 0) bits |= (uint32_t)((cw[q][nz] >> ((Y & 7) * 8)) & 0xFFull) << (8 * nz);
          rows[r] = bits;
        }
      }
    }
    __syncthreads();
    const uint32_t win = (2u << (2 * R)) - 1u;  // 2R+1 bits
    for (int i = tid; i < span * span * 8; i += 256) {
      const int zz = i & 7, rc = i >> 3, X = lo + rc / span, Y = lo + rc % span;
      const uint32_t b = (rows[X * 24 + Y] >> (8 + zz - R)) & win;
      int d = 255;
      if (b) {
        const uint32_t below = b & ((2u << R) - 1u);  // positions 0..R (at or below z)
        const uint32_t above = b >> R;                // positions R..2R (at or above z)
        int dz = 64;
        if (below) dz = R - (31 - __clz(below));
        if (above) {
          const int h = __ffs(above) - 1;
          dz = h < dz ? h : dz;
        }
        d = dz * dz;
      }
      dz2[(X * 24 + Y) * 8 + zz] = (uint8_t)d;
    }
    __syncthreads();
    for (int i = tid; i < span * 64; i += 256) {
      const int X = lo + (i >> 6), yy = (i >> 3) & 7, zz = i & 7;
      int m = 255;
      for (int dy = -R; dy <= R; ++dy) {
        const int v = dz2[(X * 24 + 8 + yy + dy) * 8 + zz];
        if (v != 255) m = min(m, v + dy * dy);
      }
      dyz2[(X * 8 + yy) * 8 + zz] = (uint8_t)m;
    }
    __syncthreads();
    bool any_left = false;
    unsigned long long n_active = 0;
    for (int x = wave; x < 8; x += 4) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      const float old = L.values[x * 64 + lane];
      bool kept = false;
      float nv = old;
      if (act) {
        int best = 0x7fffffff;
        for (int dx = -R; dx <= R; ++dx) {
          const int v = dyz2[((8 + x + dx) * 8 + y) * 8 + z];
          if (v != 255) best = min(best, v + dx * dx);
        }
        if (best != 0x7fffffff) {
          const float sign = old >= 0.0f ? 1.0f : -1.0f;
          const float v = sign * sqrtf((float)best) * vs;
          if (fabsf(v) < band) {
            kept = true;
            nv = v;
          }
        }
        if (!kept) nv = g.background;
      }
      O.values[x * 64 + lane] = nv;
      const uint64_t word = __ballot(act ? kept : false);
      if (lane == 0) O.mask[x] = word;
      any_left |= word != 0;
      n_active += __popcll(word);
    }
    if (lane == 0) {
      s_left[wave] = any_left ? 1 : 0;
      s_cnt[wave] = n_active;
    }
    __syncthreads();
    if (tid == 0) {
      keep[li] = (s_left[0] | s_left[1] | s_left[2] | s_left[3]) ? 1 : 0;
      const unsigned long long c = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
      if (active_out && c) atomicAdd(active_out, c);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ trim
// src/tools/level_set_track.rs:53-68: deactivate |v| >= hw * vs.
__global__ __launch_bounds__(64) void k_trim(GridView g, NanoLeaf* out, int* keep, float threshold) {
  const int lane = threadIdx.x;
  const int n = *g.count;
  int li_end, li_step;
  for (int li = xcd_first(n, &li_end, &li_step); li < li_end; li += li_step) {
    const NanoLeaf& L = g.leaves[li];
    NanoLeaf& O = out[li];
    if (lane < 3) O.origin[lane] = L.origin[lane];
    if (lane == 3) O.pad = 0;
    bool any_left = false;
    for (int x = 0; x < 8; ++x) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      const float v = L.values[x * 64 + lane];
      const bool kill = act && fabsf(v) >= threshold;
      O.values[x * 64 + lane] = kill ? g.background : v;
      const uint64_t word = __ballot(act && !kill);
      if (lane == 0) O.mask[x] = word;
      any_left |= word != 0;
    }
    if (lane == 0) keep[li] = any_left ? 1 : 0;
  }
}

// Flag leaves with at least one active voxel.
__global__ void k_nonempty(const NanoLeaf* leaves, const int* count, int* keep) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint64_t m = 0;
    for (int w = 0; w < 8; ++w) m |= leaves[i].mask[w];
    keep[i] = m ? 1 : 0;
  }
}

// Stable compaction of leaves: dst[pos[i]] = src[i] where keep[i].
// child_fill (optional): the leaf slots of the internal nodes, reset to -1
// for the lookup rebuild that follows.
__global__ __launch_bounds__(256) void k_compact_leaves(const NanoLeaf* src, const int* count, const int* keep,
                                                        const int* pos, NanoLeaf* dst, int* child_fill,
                                                        size_t nchild) {
  if (child_fill)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nchild; i += (size_t)gridDim.x * blockDim.x)
      child_fill[i] = -1;
  const int n = *count;
  const int sub = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = blockIdx.x * 4 + sub; i < n; i += gridDim.x * 4) {
    if (!keep[i] || pos[i] >= n) continue;
    const uint4* s = reinterpret_cast<const uint4*>(&src[i]);
    uint4* d = reinterpret_cast<uint4*>(&dst[pos[i]]);
    for (int q = lane; q < VR_LEAF_BYTES / 16; q += 64) d[q] = s[q];
  }
}

// In-place compaction with the lookup arrays updated alongside (no tree
// lookups): leaf i of the old live range [0, *count_old) with keep[i] goes
// to j = pos[i] (records from src to dst), its origin to origins[j], its
// internal-node slot to j, and its neighbour list is remapped from the old
// table (a neighbour that is kept becomes its new index, one that is
// removed becomes -1); a removed leaf's slot becomes -1.  The new count
// (*total) goes to count_out (and count_out2); count_old, nbr_old and the
// outputs are distinct arrays.
__global__ __launch_bounds__(256) void k_compact_topo(const NanoLeaf* src, const int* count_old, const int* keep,
                                                      const int* pos, const int* total, NanoLeaf* dst,
                                                      const int4* roots, int root_count, int* child, int4* origins,
                                                      const int* nbr_old, int* nbr_new, int* count_out,
                                                      int* count_out2) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int c = *total;
    *count_out = c;
    if (count_out2) *count_out2 = c;
  }
  const int n = *count_old;
  const int sub = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = blockIdx.x * 4 + sub; i < n; i += gridDim.x * 4) {
    const bool kp = keep[i] != 0;
    const int j = kp ? pos[i] : 0;
    if (kp) {
      const uint4* s = reinterpret_cast<const uint4*>(&src[i]);
      uint4* d = reinterpret_cast<uint4*>(&dst[j]);
      for (int q = lane; q < VR_LEAF_BYTES / 16; q += 64) d[q] = s[q];
      if (lane < VR_NBR) {
        const int i2 = nbr_old[(size_t)i * VR_NBR + lane];
        nbr_new[(size_t)j * VR_NBR + lane] = (i2 >= 0 && keep[i2]) ? pos[i2] : -1;
      }
    }
    if (lane == 63) {
      const int x = src[i].origin[0], y = src[i].origin[1], z = src[i].origin[2];
      const int kx = x & ~127, ky = y & ~127, kz = z & ~127;
      int lo = 0, hi = root_count;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int4 r = roots[mid];
        if (coord_less(r.x, r.y, r.z, kx, ky, kz)) lo = mid + 1; else hi = mid;
      }
      const int slot = (((x - kx) >> 3) << 8) | (((y - ky) >> 3) << 4) | ((z - kz) >> 3);
      child[(size_t)lo * 4096 + slot] = kp ? j : -1;
      if (kp) origins[j] = make_int4(x, y, z, 0);
    }
  }
}

__global__ void k_copy_int(const int* src, int* dst, int* dst2) {
  const int v = *src;
  *dst = v;
  if (dst2) *dst2 = v;
}

// p[i] = 0 for *count <= i < n (flags past a device-side count, before a
// scan whose length is a host upper bound).
__global__ void k_zero_tail(int* p, const int* count, int n) {
  for (int i = *count + (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------------ crossing list
__global__ void k_crossing_coords(const int4* origins, const int* count, const uint64_t* cross, const int* pos,
                                  int* xyz) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * 8; i += gridDim.x * blockDim.x) {
    const int li = i >> 3, x = i & 7;
    uint64_t w = cross[i];
    int p = pos[i];
    const int4 o = origins[li];
    while (w) {
      const int b = __ffsll((unsigned long long)w) - 1;
      w &= w - 1;
      xyz[3 * p] = o.x + x;
      xyz[3 * p + 1] = o.y + (b >> 3);
      xyz[3 * p + 2] = o.z + (b & 7);
      ++p;
    }
  }
}

__global__ void k_popcount_words(const uint64_t* w, const int* count, int* out) {
  const int n = *count * 8;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = __popcll(w[i]);
}

// ------------------------------------------------------------------ launch wrappers
static inline int grid_for(size_t n, int block) {
  size_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}
static inline int leaf_blocks(int cap) { return cap < 1 ? 1 : (cap < 65536 ? cap : 65536); }

hipError_t vr_fill_int(int* p, int v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t vr_build_child(const NanoLeaf* leaves, const int* count, int cap, const int4* roots, int root_count,
                          int* child, int4* origins, hipStream_t s) {
  hipLaunchKernelGGL(k_build_child, dim3(grid_for(cap, 256)), dim3(256), 0, s, leaves, count, roots, root_count, child,
                     origins);
  return hipGetLastError();
}
hipError_t vr_build_nbr(const GridView& g, int cap, int* nbr, hipStream_t s) {
  hipLaunchKernelGGL(k_build_nbr, dim3(grid_for((size_t)cap * VR_NBR, 256)), dim3(256), 0, s, g, nbr);
  return hipGetLastError();
}
size_t vr_scan_scratch_ints(size_t n) { return (n + 1023) / 1024 + 1; }
hipError_t vr_exclusive_scan(const int* in, int* out, size_t n, int* scratch, int* total, hipStream_t s,
                             const int* n_dev, int n_mult) {
  const int nb = (int)((n + 1023) / 1024);
  if (nb == 0) return total ? vr_fill_int(total, 0, 1, s) : hipSuccess;
  if (n <= VR_SCAN_SINGLE_MAX) {
    hipLaunchKernelGGL(k_scan_single<1>, dim3(1), dim3(1024), 0, s, in, out, total, (const int*)nullptr,
                       (int*)nullptr, (int*)nullptr, (int)n, n_dev, n_mult);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, in, out, scratch, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, scratch, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, s, out, scratch, n);
  return hipGetLastError();
}
hipError_t vr_exclusive_scan2(const int* in0, int* out0, int* total0, const int* in1, int* out1, int* total1,
                              size_t n, int* scratch, hipStream_t s, const int* n_dev) {
  if (n > 0 && n <= VR_SCAN_SINGLE_MAX) {
    hipLaunchKernelGGL(k_scan_single<2>, dim3(1), dim3(1024), 0, s, in0, out0, total0, in1, out1, total1, (int)n,
                       n_dev, 1);
    return hipGetLastError();
  }
  hipError_t e = vr_exclusive_scan(in0, out0, n, scratch, total0, s);
  return e != hipSuccess ? e : vr_exclusive_scan(in1, out1, n, scratch, total1, s);
}
hipError_t vr_launch_crossings(const GridView& g, int cap, uint64_t* cross, int* any, int gen, hipStream_t s) {
  hipLaunchKernelGGL(k_crossings, dim3(leaf_blocks(cap)), dim3(256), 0, s, g, cross, any, gen);
  return hipGetLastError();
}
hipError_t vr_launch_rebuild(const GridView& g, int cap, const uint64_t* cross, const int* any, int gen, NanoLeaf* out,
                             int* keep, int keep_len, int R, float vs, float band, unsigned long long* active_out,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_rebuild, dim3(leaf_blocks(cap)), dim3(256), 0, s, g, cross, any, gen, out, keep, keep_len, R,
                     vs, band, active_out);
  return hipGetLastError();
}
hipError_t vr_launch_trim(const GridView& g, int cap, NanoLeaf* out, int* keep, float threshold, hipStream_t s) {
  hipLaunchKernelGGL(k_trim, dim3(leaf_blocks(cap)), dim3(64), 0, s, g, out, keep, threshold);
  return hipGetLastError();
}
hipError_t vr_launch_nonempty(const NanoLeaf* leaves, const int* count, int cap, int* keep, hipStream_t s) {
  hipLaunchKernelGGL(k_nonempty, dim3(grid_for(cap, 256)), dim3(256), 0, s, leaves, count, keep);
  return hipGetLastError();
}
hipError_t vr_launch_compact(const NanoLeaf* src, const int* count, int cap, const int* keep, const int* pos,
                             NanoLeaf* dst, hipStream_t s, int* child_fill, size_t nchild) {
  hipLaunchKernelGGL(k_compact_leaves, dim3(grid_for(cap, 4)), dim3(256), 0, s, src, count, keep, pos, dst,
                     child_fill, nchild);
  return hipGetLastError();
}
hipError_t vr_launch_compact_topo(const NanoLeaf* src, const int* count_old, int cap, const int* keep, const int* pos,
                                  const int* total, NanoLeaf* dst, const int4* roots, int root_count, int* child,
                                  int4* origins, const int* nbr_old, int* nbr_new, int* count_out, int* count_out2,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_compact_topo, dim3(grid_for(cap, 4)), dim3(256), 0, s, src, count_old, keep, pos, total, dst,
                     roots, root_count, child, origins, nbr_old, nbr_new, count_out, count_out2);
  return hipGetLastError();
}
hipError_t vr_copy_int(const int* src, int* dst, hipStream_t s, int* dst2) {
  hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, s, src, dst, dst2);
  return hipGetLastError();
}
hipError_t vr_zero_tail(int* p, const int* count, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_tail, dim3(grid_for((size_t)(n < 1 ? 1 : n), 256)), dim3(256), 0, s, p, count, n);
  return hipGetLastError();
}
hipError_t vr_launch_popcount_words(const uint64_t* w, const int* count, int cap, int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_popcount_words, dim3(grid_for((size_t)cap * 8, 256)), dim3(256), 0, s, w, count, out);
  return hipGetLastError();
}
hipError_t vr_launch_crossing_coords(const int4* origins, const int* count, int cap, const uint64_t* cross,
                                     const int* pos, int* xyz, hipStream_t s) {
  hipLaunchKernelGGL(k_crossing_coords, dim3(grid_for((size_t)cap * 8, 256)), dim3(256), 0, s, origins, count, cross,
                     pos, xyz);
  return hipGetLastError();
}
__global__ void k_leaf_popcount(const NanoLeaf* leaves, const int* count, int* out) {
  const int n = *count * 8;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = __popcll(leaves[i >> 3].mask[i & 7]);
}
hipError_t vr_launch_leaf_popcount(const NanoLeaf* leaves, const int* count, int cap, int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_popcount, dim3(grid_for((size_t)cap * 8, 256)), dim3(256), 0, s, leaves, count, out);
  return hipGetLastError();
}
