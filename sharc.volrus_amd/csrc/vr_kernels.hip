// gfx950 kernels for the narrow-band level-set path.
//
// One wave64 per 8^3 leaf (or per marching-cubes tile): lane = y * 8 + z and
// each lane walks the 8 x-slabs, so the active mask word for slab x is the
// wave's ballot layout (mask bit = x * 64 + lane, src/tree/leaf.rs:52-56).
// Leaf bricks plus halo are staged in LDS; every count and offset is made by
// wave/block prefix scans, so outputs are deterministic.
#include <cstdlib>

#include "vr_kernels.h"

// ------------------------------------------------------------------ MC tables (device)
__constant__ uint16_t c_mc_edge_mask[256];
__constant__ uint8_t c_mc_ntri[256];
__constant__ int8_t c_mc_tri[256][16];
__constant__ int8_t c_mc_corner[8][3];
__constant__ int8_t c_mc_edge[12][2];
// For local edge e and candidate j (0..3): offset (dx, dy, dz) of a cell that
// contains the same lattice edge, and that edge's local number there.
__constant__ int8_t c_mc_cand[12][4][4];

hipError_t vr_upload_mc_tables(const uint16_t* edge_mask, const uint8_t* ntri, const int8_t* tri,
                               const int8_t* corner, const int8_t* edge, const int8_t* cand) {
  hipError_t e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_edge_mask), edge_mask, 256 * sizeof(uint16_t))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_ntri), ntri, 256)) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), tri, 256 * 16)) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_corner), corner, 24)) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_edge), edge, 24)) != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_mc_cand), cand, 12 * 4 * 4);
}

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
// New value of the active voxel (ox + x, oy + y, oz + z): back-trace, then
// trilinear sampling of the input.  Departure-cell corners inside the LDS
// brick (leaf plus one-voxel halo, 10^3) come from LDS; any other corner goes
// through the root/internal lookup, so any dt is handled.
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

// Slab variant: lane = (y, z), the wave walks the 8 x-slabs; inactive lanes
// idle while active ones compute.
template <int VK, int SCHEME>
__global__ __launch_bounds__(64) void k_advect_slab(GridView in, VelView vel, NanoLeaf* out, double dt, double tf0,
                                                    double tf1, double tf2) {
  __shared__ float brick[1000];
  __shared__ int nb[VR_NBR];
  const int lane = threadIdx.x;
  const int n = *in.count;
  const double tf[3] = {tf0, tf1, tf2};
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = in.leaves[li];
    if (lane < VR_NBR) nb[lane] = in.nbr[(size_t)li * VR_NBR + lane];
    __syncthreads();
    stage_brick_dev(in, nb, brick, lane);
    __syncthreads();
    const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
    const int y = lane >> 3, z = lane & 7;
    NanoLeaf& O = out[li];
    if (lane < 3) O.origin[lane] = L.origin[lane];
    if (lane == 3) O.pad = 0;
    if (lane < 8) O.mask[lane] = L.mask[lane];
    for (int x = 0; x < 8; ++x) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      float res = in.background;
      if (act) res = advect_voxel<VK, SCHEME>(in, vel, brick, ox, oy, oz, x, y, z, dt, tf);
      O.values[x * 64 + lane] = res;
    }
    __syncthreads();
  }
}

// Compacted variant: the active voxels of the leaf are listed in LDS by a
// ballot/popcount prefix over the 8 mask words, the wave computes them
// densely (about 3 passes of 64 for a band leaf instead of 8 slabs), and the
// leaf is then written back slab by slab, coalesced, with the background in
// the inactive positions.
template <int VK, int SCHEME>
__global__ __launch_bounds__(64) void k_advect(GridView in, VelView vel, NanoLeaf* out, double dt, double tf0,
                                               double tf1, double tf2) {
  __shared__ float brick[1000];
  __shared__ float res[512];
  __shared__ uint16_t list[512];
  __shared__ int nb[VR_NBR];
  const int lane = threadIdx.x;
  const int n = *in.count;
  const double tf[3] = {tf0, tf1, tf2};
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = in.leaves[li];
    if (lane < VR_NBR) nb[lane] = in.nbr[(size_t)li * VR_NBR + lane];
    uint64_t m[8];
    int total = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      m[x] = L.mask[x];
      if ((m[x] >> lane) & 1ull) list[total + __popcll(m[x] & below)] = (uint16_t)(x * 64 + lane);
      total += __popcll(m[x]);
    }
    __syncthreads();
    stage_brick_dev(in, nb, brick, lane);
    __syncthreads();
    const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
    for (int i = lane; i < total; i += 64) {
      const int off = list[i];
      res[off] = advect_voxel<VK, SCHEME>(in, vel, brick, ox, oy, oz, off >> 6, (off >> 3) & 7, off & 7, dt, tf);
    }
    __syncthreads();
    NanoLeaf& O = out[li];
    if (lane < 3) O.origin[lane] = L.origin[lane];
    if (lane == 3) O.pad = 0;
    if (lane < 8) O.mask[lane] = m[lane];
#pragma unroll
    for (int x = 0; x < 8; ++x)
      O.values[x * 64 + lane] = ((m[x] >> lane) & 1ull) ? res[x * 64 + lane] : in.background;
    __syncthreads();
  }
}

template <int VK, int SCHEME>
static hipError_t launch_advect_t(const GridView& in, const VelView& vel, NanoLeaf* out, int cap, double dt,
                                  const double tf[3], hipStream_t s) {
  const int blocks = cap < 1 ? 1 : (cap < 65536 ? cap : 65536);
  static const int variant = [] {
    const char* e = getenv("VOLRUS_ADVECT_VARIANT");
    return (e && e[0] == 's') ? 1 : 0;
  }();
  if (variant == 1)
    hipLaunchKernelGGL((k_advect_slab<VK, SCHEME>), dim3(blocks), dim3(64), 0, s, in, vel, out, dt, tf[0], tf[1],
                       tf[2]);
  else
    hipLaunchKernelGGL((k_advect<VK, SCHEME>), dim3(blocks), dim3(64), 0, s, in, vel, out, dt, tf[0], tf[1], tf[2]);
  return hipGetLastError();
}

hipError_t vr_launch_advect(const GridView& in, const VelView& vel, NanoLeaf* out, int cap, double dt,
                            const double tf[3], int scheme, hipStream_t s) {
#define VR_ADV(VK)                                                                         \
  case VK:                                                                                 \
    return scheme == 3 ? launch_advect_t<VK, 3>(in, vel, out, cap, dt, tf, s)               \
                       : launch_advect_t<VK, 1>(in, vel, out, cap, dt, tf, s);
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
// an active 6-neighbour whose f32 product with it is < 0.
__global__ __launch_bounds__(64) void k_crossings(GridView g, uint64_t* cross, int* any) {
  __shared__ int nb[VR_NBR];
  const int lane = threadIdx.x;
  const int n = *g.count;
  const int y = lane >> 3, z = lane & 7;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = g.leaves[li];
    if (lane < VR_NBR) nb[lane] = g.nbr[(size_t)li * VR_NBR + lane];
    __syncthreads();
    bool found = false;
    for (int x = 0; x < 8; ++x) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      bool cr = false;
      if (act) {
        const float v = L.values[x * 64 + lane];
        if (v == 0.0f) {
          cr = true;
        } else {
          const int dd[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            const int nx = x + dd[q][0], ny = y + dd[q][1], nz = z + dd[q][2];
            const int ex = nx < 0 ? -1 : (nx > 7 ? 1 : 0), ey = ny < 0 ? -1 : (ny > 7 ? 1 : 0),
                      ez = nz < 0 ? -1 : (nz > 7 ? 1 : 0);
            const int src = (ex | ey | ez) ? nb[(ex + 1) * 9 + (ey + 1) * 3 + (ez + 1)] : li;
            if (src < 0) continue;
            const int off = voxel_offset(nx, ny, nz);
            if (!((g.leaves[src].mask[off >> 6] >> (off & 63)) & 1ull)) continue;
            if (v * g.leaves[src].values[off] < 0.0f) {
              cr = true;
              break;
            }
          }
        }
      }
      const uint64_t word = __ballot(cr);
      if (lane == 0) cross[(size_t)li * 8 + x] = word;
      found |= word != 0;
    }
    if (found && lane == 0) atomicOr(any, 1);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ band rebuild
// Bounded-radius form of src/tools/level_set_track.rs:80-114: the crossing
// masks of the 27 neighbouring leaves are staged as 24 x 24 rows of 24 z-bits;
// each active voxel takes the smallest squared lattice distance to a set bit
// within |d| <= R = ceil(half_width) <= 8, which gives the brute-force result
// exactly (kept set and values).  No crossings anywhere: the grid is copied
// unchanged (the reference returns early, :85-87).
__global__ __launch_bounds__(64) void k_rebuild(GridView g, const uint64_t* cross, const int* any, NanoLeaf* out,
                                                int* keep, int R, float vs, float band) {
  __shared__ uint32_t rows[24 * 24];
  __shared__ int nb[VR_NBR];
  const int lane = threadIdx.x;
  const int n = *g.count;
  const bool have = *any != 0;
  const int y = lane >> 3, z = lane & 7;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = g.leaves[li];
    NanoLeaf& O = out[li];
    if (lane < 3) O.origin[lane] = L.origin[lane];
    if (lane == 3) O.pad = 0;
    if (!have) {
      if (lane < 8) O.mask[lane] = L.mask[lane];
      for (int x = 0; x < 8; ++x) O.values[x * 64 + lane] = L.values[x * 64 + lane];
      if (lane == 0) keep[li] = 1;
      continue;
    }
    if (lane < VR_NBR) nb[lane] = g.nbr[(size_t)li * VR_NBR + lane];
    __syncthreads();
    for (int r = lane; r < 576; r += 64) {
      const int X = r / 24, Y = r % 24;
      const int nx = X >> 3, ny = Y >> 3;
      uint32_t bits = 0;
#pragma unroll
      for (int nz = 0; nz < 3; ++nz) {
        const int src = nb[nx * 9 + ny * 3 + nz];
        if (src >= 0) bits |= (uint32_t)((cross[(size_t)src * 8 + (X & 7)] >> ((Y & 7) * 8)) & 0xFFull) << (8 * nz);
      }
      rows[r] = bits;
    }
    __syncthreads();
    bool any_left = false;
    const uint32_t win = (2u << (2 * R)) - 1u;  // 2R+1 bits
    for (int x = 0; x < 8; ++x) {
      const bool act = (L.mask[x] >> lane) & 1ull;
      const float old = L.values[x * 64 + lane];
      bool kept = false;
      float nv = old;
      if (act) {
        int best = 0x7fffffff;
        const int X = 8 + x, Y = 8 + y, Z = 8 + z;
        for (int dx = -R; dx <= R; ++dx) {
          const int dx2 = dx * dx;
          if (dx2 >= best) continue;
          for (int dy = -R; dy <= R; ++dy) {
            const int dxy = dx2 + dy * dy;
            if (dxy >= best) continue;
            const uint32_t b = (rows[(X + dx) * 24 + (Y + dy)] >> (Z - R)) & win;
            if (!b) continue;
            const uint32_t lo = b & ((2u << R) - 1u);  // positions 0..R (at or below z)
            const uint32_t hi = b >> R;                // positions R..2R (at or above z)
            int dz = 64;
            if (lo) dz = R - (31 - __clz(lo));
            if (hi) {
              const int h = __ffs(hi) - 1;
              dz = h < dz ? h : dz;
            }
            const int d2 = dxy + dz * dz;
            best = d2 < best ? d2 : best;
          }
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
    }
    if (lane == 0) keep[li] = any_left ? 1 : 0;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ trim
// src/tools/level_set_track.rs:53-68: deactivate |v| >= hw * vs.
__global__ __launch_bounds__(64) void k_trim(GridView g, NanoLeaf* out, int* keep, float threshold) {
  const int lane = threadIdx.x;
  const int n = *g.count;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
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
__global__ __launch_bounds__(256) void k_compact_leaves(const NanoLeaf* src, const int* count, const int* keep,
                                                        const int* pos, NanoLeaf* dst) {
  const int n = *count;
  const int sub = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = blockIdx.x * 4 + sub; i < n; i += gridDim.x * 4) {
    if (!keep[i] || pos[i] >= n) continue;
    const uint4* s = reinterpret_cast<const uint4*>(&src[i]);
    uint4* d = reinterpret_cast<uint4*>(&dst[pos[i]]);
    for (int q = lane; q < VR_LEAF_BYTES / 16; q += 64) d[q] = s[q];
  }
}

__global__ void k_copy_int(const int* src, int* dst) { *dst = *src; }

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

// ------------------------------------------------------------------ marching cubes
// Tile set (src/tools/volume_to_mesh.rs:57-74): leaf origins and their six
// face neighbours, sorted, de-duplicated.  Built as a merge of 7 sorted
// lists: list 0 = the leaves, list k = leaves + D[k].  An element of list k
// is kept when it is not a leaf and not in an earlier list; its rank in the
// union is the sum over lists j of the kept elements of list j below it.
__constant__ int c_dir7[7][3] = {{0, 0, 0}, {-8, 0, 0}, {8, 0, 0}, {0, -8, 0}, {0, 8, 0}, {0, 0, -8}, {0, 0, 8}};

__device__ __forceinline__ int lower_bound_origin(const int4* o, int n, int x, int y, int z) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int4 r = o[mid];
    if (coord_less(r.x, r.y, r.z, x, y, z)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_mc_keep(GridView g, int* keep) {
  const int n = *g.count;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 7 * n; t += gridDim.x * blockDim.x) {
    const int k = t / n, i = t % n;
    if (k == 0) {
      keep[t] = 1;
      continue;
    }
    const int4 o = g.origins[i];
    const int cx = o.x + c_dir7[k][0], cy = o.y + c_dir7[k][1], cz = o.z + c_dir7[k][2];
    int kp = find_leaf(g, cx, cy, cz) < 0;
    for (int j = 1; j < k && kp; ++j)
      if (find_leaf(g, cx - c_dir7[j][0], cy - c_dir7[j][1], cz - c_dir7[j][2]) >= 0) kp = 0;
    keep[t] = kp;
  }
}

__global__ void k_mc_scatter_tiles(GridView g, const int* keep, const int* pos, const int* ntiles, int4* tiles) {
  const int n = *g.count;
  const int nt = *ntiles;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 7 * n; t += gridDim.x * blockDim.x) {
    if (!keep[t]) continue;
    const int k = t / n, i = t % n;
    const int4 o = g.origins[i];
    const int ex = o.x + c_dir7[k][0], ey = o.y + c_dir7[k][1], ez = o.z + c_dir7[k][2];
    int rank = 0;
    for (int j = 0; j < 7; ++j) {
      const int lb = lower_bound_origin(g.origins, n, ex - c_dir7[j][0], ey - c_dir7[j][1], ez - c_dir7[j][2]);
      rank += (lb < n ? pos[j * n + lb] : pos[j * n + n - 1] + keep[j * n + n - 1]) - pos[j * n];
    }
    if (rank < nt) tiles[rank] = make_int4(ex, ey, ez, 0);
  }
}

// Per tile: the 27 neighbouring leaves and the 27 neighbouring tiles.
__global__ void k_mc_tile_nbr(GridView g, const int4* tiles, const int* ntiles, int* tleaf, int* ttile) {
  const int nt = *ntiles;
  const size_t total = (size_t)nt * VR_NBR;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(q / VR_NBR), d = (int)(q % VR_NBR);
    const int4 o = tiles[t];
    const int x = o.x + 8 * (d / 9 - 1), y = o.y + 8 * ((d / 3) % 3 - 1), z = o.z + 8 * (d % 3 - 1);
    tleaf[q] = find_leaf(g, x, y, z);
    const int lb = lower_bound_origin(tiles, nt, x, y, z);
    ttile[q] = (lb < nt && tiles[lb].x == x && tiles[lb].y == y && tiles[lb].z == z) ? lb : -1;
  }
}

// Owner of the lattice edge carried by local edge e of the cell at local
// (lx, ly, lz): the containing cell with the smallest visiting key
// (tile origin, lx, ly, lz) among the cells of processed tiles.  Returns the
// candidate index j (0..3).
__device__ __forceinline__ int mc_owner(int e, int lx, int ly, int lz, uint32_t processed) {
  int best = -1;
  int bk[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cx = lx + c_mc_cand[e][j][0], cy = ly + c_mc_cand[e][j][1], cz = lz + c_mc_cand[e][j][2];
    const int tx = cx >> 3, ty = cy >> 3, tz = cz >> 3;  // tile offset -1, 0 or 1
    if (!((processed >> ((tx + 1) * 9 + (ty + 1) * 3 + (tz + 1))) & 1u)) continue;
    const int k[6] = {tx, ty, tz, cx & 7, cy & 7, cz & 7};
    bool less = best < 0;
    if (!less) {
      int a = 0;
      while (a < 6 && k[a] == bk[a]) ++a;
      less = a < 6 && k[a] < bk[a];
    }
    if (less) {
      best = j;
#pragma unroll
      for (int a = 0; a < 6; ++a) bk[a] = k[a];
    }
  }
  return best;
}

__device__ __forceinline__ int mc_self_cand(int e) {
  for (int j = 0; j < 4; ++j)
    if (c_mc_cand[e][j][0] == 0 && c_mc_cand[e][j][1] == 0 && c_mc_cand[e][j][2] == 0) return j;
  return 0;
}

// Stage the 9^3 corner values of a tile (local 0..8) into LDS.
__device__ __forceinline__ void mc_stage(const GridView& g, const int* tleaf_t, float* brick) {
  for (int i = threadIdx.x; i < 729; i += 64) {
    const int bx = i / 81, by = (i / 9) % 9, bz = i % 9;
    const int d = ((bx >> 3) + 1) * 9 + ((by >> 3) + 1) * 3 + ((bz >> 3) + 1);
    const int src = tleaf_t[d];
    brick[i] = src < 0 ? g.background : g.leaves[src].values[voxel_offset(bx, by, bz)];
  }
}

// Pass 1: per cell case, owned crossing edges and triangle count; per-cell
// meta = (vertex prefix within the tile << 12) | owned-edge mask for cells
// with crossing edges; per-tile vertex and triangle totals.
__global__ __launch_bounds__(64) void k_mc_count(GridView g, const int4* tiles, const int* ntiles, const int* tleaf,
                                                 const int* ttile, float iso, uint32_t* meta, int* tv, int* tt) {
  __shared__ float brick[729];
  __shared__ int s_tleaf[VR_NBR];
  const int lane = threadIdx.x;
  const int nt = *ntiles;
  const int y = lane >> 3, z = lane & 7;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    if (lane < VR_NBR) s_tleaf[lane] = tleaf[(size_t)t * VR_NBR + lane];
    uint32_t processed = 0;
    for (int d = 0; d < VR_NBR; ++d) processed |= (ttile[(size_t)t * VR_NBR + d] >= 0 ? 1u : 0u) << d;
    __syncthreads();
    mc_stage(g, s_tleaf, brick);
    __syncthreads();
    int vcarry = 0, tcarry = 0;
    for (int x = 0; x < 8; ++x) {
      uint32_t cube = 0;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float v = brick[(x + c_mc_corner[c][0]) * 81 + (y + c_mc_corner[c][1]) * 9 + (z + c_mc_corner[c][2])];
        if (v < iso) cube |= 1u << c;
      }
      const uint32_t em = c_mc_edge_mask[cube];
      uint32_t owned = 0;
      for (int e = 0; e < 12; ++e) {
        if (!((em >> e) & 1u)) continue;
        if (mc_owner(e, x, y, z, processed) == mc_self_cand(e)) owned |= 1u << e;
      }
      const int nv = __popc(owned), ntr = c_mc_ntri[cube];
      const int vin = wave_incl_scan(nv);
      const int tin = wave_incl_scan(ntr);
      if (em) meta[(size_t)t * 512 + x * 64 + lane] = ((uint32_t)(vcarry + vin - nv) << 12) | owned;
      vcarry += __shfl(vin, 63, 64);
      tcarry += __shfl(tin, 63, 64);
    }
    if (lane == 0) {
      tv[t] = vcarry;
      tt[t] = tcarry;
    }
    __syncthreads();
  }
}

// Pass 2: write owned vertices (owner's edge orientation, f32 t, f64
// position; src/tools/volume_to_mesh.rs:162-186) and triangles in
// TRI_TABLE order (:191-200) at scanned offsets.
__global__ __launch_bounds__(64) void k_mc_emit(GridView g, const int4* tiles, const int* ntiles, const int* tleaf,
                                                const int* ttile, float iso, double vs, const uint32_t* meta,
                                                const int* vbase, const int* tbase, const int* totals, double* verts,
                                                uint32_t* tris) {
  __shared__ float brick[729];
  __shared__ int s_tleaf[VR_NBR];
  __shared__ int s_ttile[VR_NBR];
  const int lane = threadIdx.x;
  const int nt = *ntiles;
  const int y = lane >> 3, z = lane & 7;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    if (lane < VR_NBR) {
      s_tleaf[lane] = tleaf[(size_t)t * VR_NBR + lane];
      s_ttile[lane] = ttile[(size_t)t * VR_NBR + lane];
    }
    __syncthreads();
    uint32_t processed = 0;
    for (int d = 0; d < VR_NBR; ++d) processed |= (s_ttile[d] >= 0 ? 1u : 0u) << d;
    mc_stage(g, s_tleaf, brick);
    __syncthreads();
    const int4 to = tiles[t];
    const int vb = vbase[t];
    const int vmax = totals[1], tmax = totals[2];
    int tcarry = tbase[t];
    for (int x = 0; x < 8; ++x) {
      uint32_t cube = 0;
      float vals[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        vals[c] = brick[(x + c_mc_corner[c][0]) * 81 + (y + c_mc_corner[c][1]) * 9 + (z + c_mc_corner[c][2])];
        if (vals[c] < iso) cube |= 1u << c;
      }
      const uint32_t em = c_mc_edge_mask[cube];
      const int ntr = c_mc_ntri[cube];
      const int tin = wave_incl_scan(ntr);
      if (em) {
        const uint32_t m = meta[(size_t)t * 512 + x * 64 + lane];
        const uint32_t owned = m & 0xFFFu;
        int vi = vb + (int)(m >> 12);
        for (int e = 0; e < 12; ++e) {
          if (!((owned >> e) & 1u)) continue;
          if (vi >= vmax) break;
          const int a = c_mc_edge[e][0], b = c_mc_edge[e][1];
          const float va = vals[a], vbv = vals[b];
          double tpar;
          if (fabsf(vbv - va) > 1e-10f)
            tpar = (double)((iso - va) / (vbv - va));
          else
            tpar = 0.5;
          const int ax = to.x + x + c_mc_corner[a][0], ay = to.y + y + c_mc_corner[a][1], az = to.z + z + c_mc_corner[a][2];
          const int bx = to.x + x + c_mc_corner[b][0], by = to.y + y + c_mc_corner[b][1], bz = to.z + z + c_mc_corner[b][2];
          const double pax = (double)ax * vs, pay = (double)ay * vs, paz = (double)az * vs;
          const double pbx = (double)bx * vs, pby = (double)by * vs, pbz = (double)bz * vs;
          verts[3 * (size_t)vi + 0] = pax + tpar * (pbx - pax);
          verts[3 * (size_t)vi + 1] = pay + tpar * (pby - pay);
          verts[3 * (size_t)vi + 2] = paz + tpar * (pbz - paz);
          ++vi;
        }
        int ev[12];
        for (int e = 0; e < 12; ++e) {
          ev[e] = 0;
          if (!((em >> e) & 1u)) continue;
          const int j = mc_owner(e, x, y, z, processed);
          if (j < 0) continue;
          const int cx = x + c_mc_cand[e][j][0], cy = y + c_mc_cand[e][j][1], cz = z + c_mc_cand[e][j][2];
          const int ep = c_mc_cand[e][j][3];
          const int d = ((cx >> 3) + 1) * 9 + ((cy >> 3) + 1) * 3 + ((cz >> 3) + 1);
          const int ot = s_ttile[d];
          if (ot < 0) continue;
          const uint32_t om = meta[(size_t)ot * 512 + ((cx & 7) << 6) + ((cy & 7) << 3) + (cz & 7)];
          ev[e] = vbase[ot] + (int)(om >> 12) + __popc((om & 0xFFFu) & ((1u << ep) - 1u));
        }
        int ti = tcarry + tin - ntr;
        for (int q = 0; q < ntr && ti < tmax; ++q) {
          tris[3 * (size_t)ti + 0] = (uint32_t)ev[c_mc_tri[cube][3 * q]];
          tris[3 * (size_t)ti + 1] = (uint32_t)ev[c_mc_tri[cube][3 * q + 1]];
          tris[3 * (size_t)ti + 2] = (uint32_t)ev[c_mc_tri[cube][3 * q + 2]];
          ++ti;
        }
      }
      tcarry += __shfl(tin, 63, 64);
    }
    __syncthreads();
  }
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
hipError_t vr_exclusive_scan(const int* in, int* out, size_t n, int* scratch, int* total, hipStream_t s) {
  const int nb = (int)((n + 1023) / 1024);
  if (nb == 0) return total ? vr_fill_int(total, 0, 1, s) : hipSuccess;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, in, out, scratch, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, scratch, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, s, out, scratch, n);
  return hipGetLastError();
}
hipError_t vr_launch_crossings(const GridView& g, int cap, uint64_t* cross, int* any, hipStream_t s) {
  hipLaunchKernelGGL(k_crossings, dim3(leaf_blocks(cap)), dim3(64), 0, s, g, cross, any);
  return hipGetLastError();
}
hipError_t vr_launch_rebuild(const GridView& g, int cap, const uint64_t* cross, const int* any, NanoLeaf* out,
                             int* keep, int R, float vs, float band, hipStream_t s) {
  hipLaunchKernelGGL(k_rebuild, dim3(leaf_blocks(cap)), dim3(64), 0, s, g, cross, any, out, keep, R, vs, band);
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
                             NanoLeaf* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_compact_leaves, dim3(grid_for(cap, 4)), dim3(256), 0, s, src, count, keep, pos, dst);
  return hipGetLastError();
}
hipError_t vr_copy_int(const int* src, int* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, s, src, dst);
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
hipError_t vr_launch_mc_keep(const GridView& g, int cap, int* keep, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_keep, dim3(grid_for((size_t)cap * 7, 256)), dim3(256), 0, s, g, keep);
  return hipGetLastError();
}
hipError_t vr_launch_mc_scatter(const GridView& g, int cap, const int* keep, const int* pos, const int* ntiles,
                                int4* tiles, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_scatter_tiles, dim3(grid_for((size_t)cap * 7, 256)), dim3(256), 0, s, g, keep, pos, ntiles,
                     tiles);
  return hipGetLastError();
}
hipError_t vr_launch_mc_tile_nbr(const GridView& g, const int4* tiles, const int* ntiles, int tcap, int* tleaf,
                                 int* ttile, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_tile_nbr, dim3(grid_for((size_t)tcap * VR_NBR, 256)), dim3(256), 0, s, g, tiles, ntiles,
                     tleaf, ttile);
  return hipGetLastError();
}
hipError_t vr_launch_mc_count(const GridView& g, const int4* tiles, const int* ntiles, int tcap, const int* tleaf,
                              const int* ttile, float iso, uint32_t* meta, int* tv, int* tt, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_count, dim3(leaf_blocks(tcap)), dim3(64), 0, s, g, tiles, ntiles, tleaf, ttile, iso, meta,
                     tv, tt);
  return hipGetLastError();
}
hipError_t vr_launch_mc_emit(const GridView& g, const int4* tiles, const int* ntiles, int tcap, const int* tleaf,
                             const int* ttile, float iso, double vs, const uint32_t* meta, const int* vbase,
                             const int* tbase, double* verts, uint32_t* tris, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_emit, dim3(leaf_blocks(tcap)), dim3(64), 0, s, g, tiles, ntiles, tleaf, ttile, iso, vs,
                     meta, vbase, tbase, ntiles, verts, tris);
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
