// Marching cubes (src/tools/volume_to_mesh.rs:46-201) over one or many
// level sets at once.
//
// Grids are concatenated: grid g owns leaves [leaf_base[g], leaf_base[g+1])
// of the global leaf numbering, tiles [tile_base[g], tile_base[g+1]) of the
// global tile list (tile.w = g), and the vertex/triangle ranges that the
// scans give those tiles.  A single grid is the case G = 1.
//
// Tile set (volume_to_mesh.rs:57-74): leaf origins and their six face
// neighbours, sorted, de-duplicated.  Built per grid as a merge of 7 sorted
// lists: list 0 = the leaves, list k = leaves + D[k].  An element of list k
// is kept when it is not a leaf and not in an earlier list; its rank in the
// grid's union is the sum over lists j of the kept elements of list j below
// it (binary searches, no sort).
//
// Numbering: the reference creates a vertex at the first cell (in tile,
// lx, ly, lz order) that meets a lattice edge.  That cell is the edge's
// owner: the processed cell with the smallest (tile origin, lx, ly, lz)
// among the four that contain the edge.  Pass 1 stores per cell the owned
// crossing edges and their prefix inside the tile; tiles are scanned; pass 2
// writes vertices in the owner's edge orientation and triangles in
// TRI_TABLE order at the scanned offsets.  The result is the reference's
// vertex and triangle lists, bit for bit.
#include "vr_internal.h"

// ------------------------------------------------------------------ tables
__constant__ uint16_t c_mc_edge_mask[256];
__constant__ uint8_t c_mc_ntri[256];
__constant__ int8_t c_mc_tri[256][16];
__constant__ int8_t c_mc_corner[8][3];
__constant__ int8_t c_mc_edge[12][2];
// For local edge e and candidate j (0..3): offset (dx, dy, dz) of a cell that
// contains the same lattice edge, and that edge's local number there.
__constant__ int8_t c_mc_cand[12][4][4];
__constant__ int c_dir7[7][3] = {{0, 0, 0}, {-8, 0, 0}, {8, 0, 0}, {0, -8, 0}, {0, 8, 0}, {0, 0, -8}, {0, 0, 8}};

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

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lower_bound_origin(const int4* o, int lo, int hi, int x, int y, int z) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int4 r = o[mid];
    if (coord_less(r.x, r.y, r.z, x, y, z)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Index g with base[g] <= i < base[g + 1] (base has G + 1 entries).
__device__ __forceinline__ int segment_of(const int* base, int G, long long i) {
  int lo = 0, hi = G;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((long long)base[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------ tile set
__global__ void k_mc_keep(const GridView* views, const int* leaf_base, int G, int* keep) {
  const long long total = 7LL * leaf_base[G];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int g = segment_of(leaf_base, G, t / 7);  // 7 * leaf_base is monotone like leaf_base
    const long long loc = t - 7LL * leaf_base[g];
    const int n = leaf_base[g + 1] - leaf_base[g];
    const int k = (int)(loc / n), i = (int)(loc % n);
    if (k == 0) {
      keep[t] = 1;
      continue;
    }
    const GridView& v = views[g];
    const int4 o = v.origins[i];
    const int cx = o.x + c_dir7[k][0], cy = o.y + c_dir7[k][1], cz = o.z + c_dir7[k][2];
    int kp = find_leaf(v, cx, cy, cz) < 0;
    for (int j = 1; j < k && kp; ++j)
      if (find_leaf(v, cx - c_dir7[j][0], cy - c_dir7[j][1], cz - c_dir7[j][2]) >= 0) kp = 0;
    keep[t] = kp;
  }
}

// tile_base[g] = first tile of grid g (G + 1 entries; tile_base[G] = total).
__global__ void k_mc_tile_base(const int* leaf_base, int G, const int* pos, const int* ntiles, int* tile_base) {
  const long long total7 = 7LL * leaf_base[G];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
    const long long s = 7LL * leaf_base[g];
    tile_base[g] = s < total7 ? pos[s] : *ntiles;
  }
}

__global__ void k_mc_scatter_tiles(const GridView* views, const int* leaf_base, int G, const int* keep, const int* pos,
                                   const int* ntiles, int4* tiles) {
  const long long total = 7LL * leaf_base[G];
  const int nt = *ntiles;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    if (!keep[t]) continue;
    const int g = segment_of(leaf_base, G, t / 7);
    const long long seg = 7LL * leaf_base[g];
    const long long loc = t - seg;
    const int n = leaf_base[g + 1] - leaf_base[g];
    const int k = (int)(loc / n), i = (int)(loc % n);
    const GridView& v = views[g];
    const int4 o = v.origins[i];
    const int ex = o.x + c_dir7[k][0], ey = o.y + c_dir7[k][1], ez = o.z + c_dir7[k][2];
    long long rank = pos[seg];
    for (int j = 0; j < 7; ++j) {
      const long long lj = seg + (long long)j * n;
      const int lb = lower_bound_origin(v.origins, 0, n, ex - c_dir7[j][0], ey - c_dir7[j][1], ez - c_dir7[j][2]);
      rank += (lb < n ? pos[lj + lb] : pos[lj + n - 1] + keep[lj + n - 1]) - pos[lj];
    }
    if (rank < nt) tiles[rank] = make_int4(ex, ey, ez, g);
  }
}

// Per tile: the 27 neighbouring leaves and the 27 neighbouring tiles (-1 when
// that tile is not processed).
__global__ void k_mc_tile_nbr(const GridView* views, const int4* tiles, const int* ntiles, const int* tile_base,
                              int* tleaf, int* ttile) {
  const int nt = *ntiles;
  const size_t total = (size_t)nt * VR_NBR;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(q / VR_NBR), d = (int)(q % VR_NBR);
    const int4 o = tiles[t];
    const int g = o.w;
    const int x = o.x + 8 * (d / 9 - 1), y = o.y + 8 * ((d / 3) % 3 - 1), z = o.z + 8 * (d % 3 - 1);
    tleaf[q] = find_leaf(views[g], x, y, z);
    const int lo = tile_base[g], hi = tile_base[g + 1];
    const int lb = lower_bound_origin(tiles, lo, hi, x, y, z);
    ttile[q] = (lb < hi && tiles[lb].x == x && tiles[lb].y == y && tiles[lb].z == z) ? lb : -1;
  }
}

// ------------------------------------------------------------------ per-cell rules
// Owner of the lattice edge carried by local edge e of the cell at local
// (lx, ly, lz): the candidate j (0..3) with the smallest visiting key
// (tile origin, lx, ly, lz) among the cells of processed tiles.
//
// The key is one integer: the neighbour-tile index (tx+1)*9 + (ty+1)*3 +
// (tz+1), whose order is the lexicographic order of tile origins around the
// cell's tile, times 512, plus the local offset lx*64 + ly*8 + lz.
__device__ __forceinline__ int mc_owner(int e, int lx, int ly, int lz, uint32_t processed) {
  int best = -1, bkey = 0x7fffffff;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cx = lx + c_mc_cand[e][j][0], cy = ly + c_mc_cand[e][j][1], cz = lz + c_mc_cand[e][j][2];
    const int t = ((cx >> 3) + 1) * 9 + ((cy >> 3) + 1) * 3 + ((cz >> 3) + 1);
    const int key = (t << 9) + ((cx & 7) << 6) + ((cy & 7) << 3) + (cz & 7);
    const bool ok = (processed >> t) & 1u;
    best = (ok && key < bkey) ? j : best;
    bkey = (ok && key < bkey) ? key : bkey;
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

// ------------------------------------------------------------------ shared per-cell pass
// Index of a cell among the 169 cells of a tile with a local coordinate 7
// (-1 for the other 343).  An owner in another tile always has one: its tile
// precedes the cell's, so its first differing tile offset is -1 and its local
// coordinate on that axis is 7.
__device__ __forceinline__ int mc_face_index(int lx, int ly, int lz) {
  if (lx == 7) return ly * 8 + lz;
  if (ly == 7) return 64 + lx * 8 + lz;
  if (lz == 7) return 120 + lx * 7 + ly;
  return -1;
}
#define MC_FACE 169

__device__ __forceinline__ uint32_t mc_case(const float* brick, int x, int y, int z, float iso, float vals[8]) {
  uint32_t cube = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    vals[c] = brick[(x + c_mc_corner[c][0]) * 81 + (y + c_mc_corner[c][1]) * 9 + (z + c_mc_corner[c][2])];
    if (vals[c] < iso) cube |= 1u << c;
  }
  return cube;
}

__device__ __forceinline__ uint32_t mc_owned_mask(uint32_t em, int x, int y, int z, uint32_t processed) {
  uint32_t owned = 0;
  for (int e = 0; e < 12; ++e) {
    if (!((em >> e) & 1u)) continue;
    if (mc_owner(e, x, y, z, processed) == mc_self_cand(e)) owned |= 1u << e;
  }
  return owned;
}

// ------------------------------------------------------------------ pass 1
// Per cell: case, owned crossing edges, triangle count.  The face cells
// (mc_face_index >= 0) with crossing edges store meta = (vertex prefix inside
// the tile << 12) | owned-edge mask for the emit pass of later tiles; per-tile
// vertex and triangle totals.
__global__ __launch_bounds__(64) void k_mc_count(const GridView* views, const int4* tiles, const int* ntiles,
                                                 const int* tleaf, const int* ttile, float iso, uint32_t* meta,
                                                 int* tv, int* tt) {
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
    const GridView& g = views[tiles[t].w];
    uint32_t processed = 0;
    for (int d = 0; d < VR_NBR; ++d) processed |= (s_ttile[d] >= 0 ? 1u : 0u) << d;
    mc_stage(g, s_tleaf, brick);
    __syncthreads();
    int vcarry = 0, tcarry = 0;
    for (int x = 0; x < 8; ++x) {
      float vals[8];
      const uint32_t cube = mc_case(brick, x, y, z, iso, vals);
      const uint32_t em = c_mc_edge_mask[cube];
      const uint32_t owned = em ? mc_owned_mask(em, x, y, z, processed) : 0u;
      const int nv = __popc(owned), ntr = c_mc_ntri[cube];
      const int vin = wave_incl_scan(nv);
      const int tin = wave_incl_scan(ntr);
      const int fi = mc_face_index(x, y, z);
      if (em && fi >= 0) meta[(size_t)t * MC_FACE + fi] = ((uint32_t)(vcarry + vin - nv) << 12) | owned;
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

// ------------------------------------------------------------------ pass 2
// Owned vertices (owner's edge orientation, f32 t, f64 position;
// volume_to_mesh.rs:162-186) and triangles in TRI_TABLE order (:191-200).
// The tile's own cell meta is rebuilt in LDS; owners in earlier tiles are
// read from their face meta.  Vertex indices are relative to the tile's grid.
__global__ __launch_bounds__(64) void k_mc_emit(const GridView* views, const int4* tiles, const int* ntiles,
                                                const int* tleaf, const int* ttile, const int* tile_base, float iso,
                                                const uint32_t* meta, const int* vbase, const int* tbase,
                                                const int* totals, double* verts, uint32_t* tris) {
  __shared__ float brick[729];
  __shared__ uint32_t s_meta[512];
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
    const int4 to = tiles[t];
    const GridView& g = views[to.w];
    const double vs = g.vs;
    const int g0 = tile_base[to.w];
    const int vrel = g0 < nt ? vbase[g0] : totals[1];  // first vertex of this grid
    uint32_t processed = 0;
    for (int d = 0; d < VR_NBR; ++d) processed |= (s_ttile[d] >= 0 ? 1u : 0u) << d;
    mc_stage(g, s_tleaf, brick);
    __syncthreads();
    {
      int vcarry = 0;
      for (int x = 0; x < 8; ++x) {
        float vals[8];
        const uint32_t em = c_mc_edge_mask[mc_case(brick, x, y, z, iso, vals)];
        const uint32_t owned = em ? mc_owned_mask(em, x, y, z, processed) : 0u;
        const int nv = __popc(owned);
        const int vin = wave_incl_scan(nv);
        s_meta[x * 64 + lane] = ((uint32_t)(vcarry + vin - nv) << 12) | owned;
        vcarry += __shfl(vin, 63, 64);
      }
    }
    __syncthreads();
    const int vb = vbase[t];
    const int vmax = totals[1], tmax = totals[2];
    int tcarry = tbase[t];
    for (int x = 0; x < 8; ++x) {
      float vals[8];
      const uint32_t cube = mc_case(brick, x, y, z, iso, vals);
      const uint32_t em = c_mc_edge_mask[cube];
      const int ntr = c_mc_ntri[cube];
      const int tin = wave_incl_scan(ntr);
      if (em) {
        const uint32_t m = s_meta[x * 64 + lane];
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
          uint32_t om;
          int obase;
          if (d == 13) {
            om = s_meta[(cx << 6) + (cy << 3) + cz];
            obase = vb;
          } else {
            const int ot = s_ttile[d];
            const int fi = mc_face_index(cx & 7, cy & 7, cz & 7);
            if (ot < 0 || fi < 0) continue;
            om = meta[(size_t)ot * MC_FACE + fi];
            obase = vbase[ot];
          }
          ev[e] = obase - vrel + (int)(om >> 12) + __popc((om & 0xFFFu) & ((1u << ep) - 1u));
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

// Per-grid vertex and triangle ranges: out[4 g] = first vertex, [4 g + 1] =
// vertex count, [4 g + 2] = first triangle, [4 g + 3] = triangle count.
__global__ void k_mc_ranges(const int* tile_base, int G, const int* ntiles, const int* vbase, const int* tbase,
                            const int* totals, int* out) {
  const int nt = *ntiles;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    const int a = tile_base[g], b = tile_base[g + 1];
    const int va = a < nt ? vbase[a] : totals[1], vb = b < nt ? vbase[b] : totals[1];
    const int ta = a < nt ? tbase[a] : totals[2], tb = b < nt ? tbase[b] : totals[2];
    out[4 * g] = va;
    out[4 * g + 1] = vb - va;
    out[4 * g + 2] = ta;
    out[4 * g + 3] = tb - ta;
  }
}

// ------------------------------------------------------------------ host driver
static int blocks_for(size_t n, int b) {
  size_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// Phase A: tile list for G grids (leaf counts known on the host: ltot).
// Leaves the tile count in s.totals[0] and tile_base; caller reads nt.
int mc_tiles_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot) {
  hipError_t e;
  const size_t n7 = (size_t)(7 * ltot);
  if ((e = s.keep.ensure(sizeof(int) * (n7 + 1))) != hipSuccess || (e = s.pos.ensure(sizeof(int) * (n7 + 1))) != hipSuccess ||
      (e = s.sums.ensure(sizeof(int) * (vr_scan_scratch_ints(n7) + 1))) != hipSuccess ||
      (e = s.totals.ensure(sizeof(int) * 4)) != hipSuccess ||
      (e = s.tile_base.ensure(sizeof(int) * (size_t)(G + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  hipLaunchKernelGGL(k_mc_keep, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views, d_leaf_base, G,
                     s.keep.as<int>());
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = vr_exclusive_scan(s.keep.as<int>(), s.pos.as<int>(), n7, s.sums.as<int>(), s.totals.as<int>(),
                             ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "mc tiles");
  hipLaunchKernelGGL(k_mc_tile_base, dim3(blocks_for(G + 1, 256)), dim3(256), 0, ctx->stream, d_leaf_base, G,
                     s.pos.as<int>(), s.totals.as<int>(), s.tile_base.as<int>());
  return (e = hipGetLastError()) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "mc tile base");
}

// Phase B (tile capacity tcap known): scatter, neighbours, count, scans.
int mc_count_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot,
                   int tcap, float iso, double algorithmic_bytes) {
  hipError_t e;
  const size_t n7 = (size_t)(7 * ltot);
  const size_t tc = (size_t)(tcap + 1);
  if ((e = s.tiles.ensure(sizeof(int4) * tc)) != hipSuccess || (e = s.tleaf.ensure(sizeof(int) * VR_NBR * tc)) != hipSuccess ||
      (e = s.ttile.ensure(sizeof(int) * VR_NBR * tc)) != hipSuccess ||
      (e = s.meta.ensure(sizeof(uint32_t) * MC_FACE * tc)) != hipSuccess || (e = s.tv.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.tt.ensure(sizeof(int) * tc)) != hipSuccess || (e = s.vb.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.tb.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.sums2.ensure(sizeof(int) * (vr_scan_scratch_ints(tc) + 1))) != hipSuccess ||
      (e = s.ranges.ensure(sizeof(int) * 4 * (size_t)(G + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  int* totals = s.totals.as<int>();
  hipLaunchKernelGGL(k_mc_scatter_tiles, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views, d_leaf_base, G,
                     s.keep.as<int>(), s.pos.as<int>(), totals, s.tiles.as<int4>());
  hipLaunchKernelGGL(k_mc_tile_nbr, dim3(blocks_for((size_t)tcap * VR_NBR, 256)), dim3(256), 0, ctx->stream, d_views,
                     s.tiles.as<int4>(), totals, s.tile_base.as<int>(), s.tleaf.as<int>(), s.ttile.as<int>());
  if ((e = hipGetLastError()) != hipSuccess) return vr_hip_fail(ctx, e, "mc tiles");
  {
    TimedLaunch tl(ctx, TC_MC, algorithmic_bytes);
    hipLaunchKernelGGL(k_mc_count, dim3(blocks_for(tcap, 1)), dim3(64), 0, ctx->stream, d_views, s.tiles.as<int4>(),
                       totals, s.tleaf.as<int>(), s.ttile.as<int>(), iso, s.meta.as<uint32_t>(), s.tv.as<int>(),
                       s.tt.as<int>());
  }
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = vr_exclusive_scan(s.tv.as<int>(), s.vb.as<int>(), (size_t)tcap, s.sums2.as<int>(), totals + 1,
                             ctx->stream)) != hipSuccess ||
      (e = vr_exclusive_scan(s.tt.as<int>(), s.tb.as<int>(), (size_t)tcap, s.sums2.as<int>(), totals + 2,
                             ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "mc count");
  hipLaunchKernelGGL(k_mc_ranges, dim3(blocks_for(G, 256)), dim3(256), 0, ctx->stream, s.tile_base.as<int>(), G, totals,
                     s.vb.as<int>(), s.tb.as<int>(), totals, s.ranges.as<int>());
  return (e = hipGetLastError()) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "mc ranges");
}

// Phase C: emit into verts/tris (capacities from the counts).
int mc_emit_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, int tcap, float iso, double* verts, uint32_t* tris,
                  double algorithmic_bytes) {
  hipError_t e;
  {
    TimedLaunch tl(ctx, TC_MC, algorithmic_bytes);
    hipLaunchKernelGGL(k_mc_emit, dim3(blocks_for(tcap, 1)), dim3(64), 0, ctx->stream, d_views, s.tiles.as<int4>(),
                       s.totals.as<int>(), s.tleaf.as<int>(), s.ttile.as<int>(), s.tile_base.as<int>(), iso,
                       s.meta.as<uint32_t>(), s.vb.as<int>(), s.tb.as<int>(), s.totals.as<int>(), verts, tris);
  }
  return (e = hipGetLastError()) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "mc emit");
}

// Single grid (volume_to_mesh) through the multi-grid passes with G = 1.
int mc_run(vr_ctx* ctx, const vr_grid* g, float iso, MCScratch* unused, vr_mesh* m) {
  (void)unused;
  hipError_t e;
  const int n = g->topo->count;
  MCMulti& s = ctx->mc;
  if ((e = s.views.ensure(sizeof(GridView))) != hipSuccess || (e = s.leaf_base.ensure(sizeof(int) * 2)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  const GridView v = make_view(g);
  const int lb[2] = {0, n};
  if ((e = hipMemcpyAsync(s.views.p, &v, sizeof(GridView), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(s.leaf_base.p, lb, sizeof(lb), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  int rc = mc_tiles_multi(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n);
  if (rc != VR_OK) return rc;
  int nt = 0;
  if ((e = hipMemcpyAsync(&nt, s.totals.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "readback");
  rc = mc_count_multi(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, nt, iso, (double)VR_LEAF_BYTES * n);
  if (rc != VR_OK) return rc;
  int th[3] = {0, 0, 0};
  if ((e = hipMemcpyAsync(th, s.totals.p, sizeof(int) * 3, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "readback");
  m->nv = (uint64_t)th[1];
  m->nt = (uint64_t)th[2];
  if (m->cap_v < m->nv + 1) {
    if (m->d_v) (void)hipFree(m->d_v);
    m->d_v = nullptr;
    m->cap_v = 0;
    if ((e = hipMalloc(&m->d_v, sizeof(double) * 3 * (size_t)(th[1] + 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc mesh");
    m->cap_v = (uint64_t)th[1] + 1;
  }
  if (m->cap_t < m->nt + 1) {
    if (m->d_t) (void)hipFree(m->d_t);
    m->d_t = nullptr;
    m->cap_t = 0;
    if ((e = hipMalloc(&m->d_t, sizeof(uint32_t) * 3 * (size_t)(th[2] + 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "hipMalloc mesh");
    m->cap_t = (uint64_t)th[2] + 1;
  }
  return mc_emit_multi(ctx, s, s.views.as<GridView>(), nt, iso, m->d_v, m->d_t, 24.0 * th[1] + 12.0 * th[2]);
}
