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
#include <mutex>

#include "vr_internal.h"

// ------------------------------------------------------------------ tables
// Per case: crossing-edge mask | triangle count << 12.
__constant__ uint16_t c_mc_emn[256];
// Per case: TRI_TABLE row, one nibble per entry (entry k at bits 4k; 15 pads).
__constant__ uint64_t c_mc_trin[256];
// For local edge e and candidate j (0..3): offset (dx, dy, dz) of a cell that
// contains the same lattice edge and that edge's local number there, packed
// as 4 int8 (dx, dy, dz, e').
__constant__ uint32_t c_mc_cand4[12 * 4];
// Per local edge, the edge's reference orientation: corner a's offset from
// the cell as packed fields (x << 8 | y << 4 | z, each 0 or 1) | corner b's
// << 16.  Added to a packed cell (mc_pack) they give the corners' packed
// lattice points (each coordinate 0..8).
__constant__ uint32_t c_mc_edgeinfo[12];
// Candidate of each local edge that is the cell itself.
__constant__ int c_mc_self[12];
// Per cell class and local edge: bit set of the neighbour tiles holding a
// candidate ordered before the first candidate in the cell's own tile.  The
// edge's owner lies in another tile when one of them is processed.
__constant__ uint32_t c_mc_fbefore[27 * 12];
// Per local edge: its slot in a tile's lattice-edge table for the cell at
// local (x, y, z) is const + x * mx + y * my + z, packed const | mx << 11 |
// my << 18.  The table holds the 1944 edges of the tile's 8^3 cells: x edges
// at (0..7, 0..8, 0..8), then y edges at (0..8, 0..7, 0..8), then z edges at
// (0..8, 0..8, 0..7), each by its lower endpoint.
__constant__ uint32_t c_mc_eslot[12];
// Per cell class (27: per axis local 0 / 1..6 / 7) and local edge: the four
// candidates in ascending visiting-key order, one byte each: j << 5 | the
// neighbour-tile index (0..26) that candidate lies in.
__constant__ uint32_t c_mc_ownsort[27 * 12];
// Per cell class and local edge: bit set of the neighbour tiles (0..26) that
// hold a candidate ordered before the cell itself.  The cell owns the edge
// when none of those tiles is processed.
__constant__ uint32_t c_mc_before[27 * 12];
__constant__ int c_dir7[7][3] = {{0, 0, 0}, {-8, 0, 0}, {8, 0, 0}, {0, -8, 0}, {0, 8, 0}, {0, 0, -8}, {0, 0, 8}};

// Uploads the tables derived from the reference's EDGE_TABLE / TRI_TABLE
// (volume_to_mesh.rs:6-44) and its corner/edge numbering.
hipError_t vr_upload_mc_tables(const uint16_t* edge_mask, const uint8_t* ntri, const int8_t* tri,
                               const int8_t* corner, const int8_t* edge, const int8_t* cand) {
  uint16_t emn[256];
  for (int c = 0; c < 256; ++c) emn[c] = (uint16_t)(edge_mask[c] | (ntri[c] << 12));
  uint32_t cand4[48], einfo[12], ownsort[27 * 12], before[27 * 12], fbefore[27 * 12], eslot[12];
  int self[12];
  for (int e = 0; e < 12; ++e) {
    self[e] = 0;
    for (int j = 0; j < 4; ++j) {
      const int8_t* q = cand + (e * 4 + j) * 4;
      uint32_t v = 0;
      for (int k = 0; k < 4; ++k) v |= (uint32_t)(uint8_t)q[k] << (8 * k);
      cand4[e * 4 + j] = v;
      if (q[0] == 0 && q[1] == 0 && q[2] == 0) self[e] = j;
    }
    const int ca = edge[2 * e], cb = edge[2 * e + 1];
    const uint32_t ab = (uint32_t)(corner[3 * ca] | (corner[3 * ca + 1] << 1) | (corner[3 * ca + 2] << 2));
    const uint32_t bb = (uint32_t)(corner[3 * cb] | (corner[3 * cb + 1] << 1) | (corner[3 * cb + 2] << 2));
    const auto pack = [](uint32_t o) { return ((o & 1u) << 8) | (((o >> 1) & 1u) << 4) | ((o >> 2) & 1u); };
    einfo[e] = pack(ab) | (pack(bb) << 16);
    const uint32_t lo = ab & bb, ax = ab ^ bb;  // lower endpoint offset bits, axis bit
    const uint32_t lx = lo & 1u, ly = (lo >> 1) & 1u, lz = lo >> 2;
    uint32_t k0, mx, my;
    if (ax == 1u) {
      k0 = lx * 81 + ly * 9 + lz;
      mx = 81;
      my = 9;
    } else if (ax == 2u) {
      k0 = 648 + lx * 72 + ly * 9 + lz;
      mx = 72;
      my = 9;
    } else {
      k0 = 1296 + lx * 72 + ly * 8 + lz;
      mx = 72;
      my = 8;
    }
    // in bytes of the 16-bit table: twice the slot base, x and y strides
    eslot[e] = (2 * k0) | ((2 * mx) << 12) | ((2 * my) << 20);
  }
  const int rep[3] = {0, 3, 7};
  for (int c = 0; c < 27; ++c)
    for (int e = 0; e < 12; ++e) {
      const int l[3] = {rep[c / 9], rep[(c / 3) % 3], rep[c % 3]};
      int key[4], tile[4], ord[4] = {0, 1, 2, 3};
      for (int j = 0; j < 4; ++j) {
        const int8_t* q = cand + (e * 4 + j) * 4;
        const int cx = l[0] + q[0], cy = l[1] + q[1], cz = l[2] + q[2];
        tile[j] = ((cx >> 3) + 1) * 9 + ((cy >> 3) + 1) * 3 + ((cz >> 3) + 1);
        key[j] = (tile[j] << 9) + ((cx & 7) << 6) + ((cy & 7) << 3) + (cz & 7);
      }
      for (int i = 1; i < 4; ++i)
        for (int k = i; k > 0 && key[ord[k]] < key[ord[k - 1]]; --k) {
          const int tmp = ord[k];
          ord[k] = ord[k - 1];
          ord[k - 1] = tmp;
        }
      uint32_t v = 0, b = 0, fb = 0;
      bool seen_self = false, seen_tile = false;
      for (int k = 0; k < 4; ++k) {
        v |= (uint32_t)((ord[k] << 5) | tile[ord[k]]) << (8 * k);
        if (ord[k] == self[e]) seen_self = true;
        if (tile[ord[k]] == 13) seen_tile = true;
        if (!seen_self) b |= 1u << tile[ord[k]];
        if (!seen_tile) fb |= 1u << tile[ord[k]];
      }
      ownsort[c * 12 + e] = v;
      before[c * 12 + e] = b;
      fbefore[c * 12 + e] = fb;
    }
  hipError_t e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_emn), emn, sizeof(emn))) != hipSuccess) return e;
  uint64_t trin[256];
  for (int c = 0; c < 256; ++c) {
    trin[c] = 0;
    for (int k = 0; k < 16; ++k) trin[c] |= (uint64_t)(tri[c * 16 + k] < 0 ? 15 : tri[c * 16 + k]) << (4 * k);
  }
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_trin), trin, sizeof(trin))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_cand4), cand4, sizeof(cand4))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_edgeinfo), einfo, sizeof(einfo))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_self), self, sizeof(self))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_before), before, sizeof(before))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_fbefore), fbefore, sizeof(fbefore))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_eslot), eslot, sizeof(eslot))) != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_mc_ownsort), ownsort, sizeof(ownsort));
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
// Entries [7 * leaf_base[G], n7cap) (a host upper bound on the length) are
// cleared for the scan.
// SINGLE: one grid given by value (sv, its leaf count on the device), whose
// view and leaf range [0, count) this launch also writes to views[0] and
// leaf_base[0..1] for the later passes; it reads neither.
template <bool SINGLE>
__global__ void k_mc_keep(const GridView* views, const int* leaf_base, int G, int* keep, long long n7cap, GridView sv,
                          GridView* views_out, int* leaf_base_out) {
  const int n1 = SINGLE ? *sv.count : 0;
  if (SINGLE && blockIdx.x == 0 && threadIdx.x == 0) {
    views_out[0] = sv;
    leaf_base_out[0] = 0;
    leaf_base_out[1] = n1;
  }
  const long long total = 7LL * (SINGLE ? n1 : leaf_base[G]);
  for (long long t = total + blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n7cap;
       t += (long long)gridDim.x * blockDim.x)
    keep[t] = 0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int g = SINGLE ? 0 : segment_of(leaf_base, G, t / 7);  // 7 * leaf_base is monotone like leaf_base
    const long long loc = SINGLE ? t : t - 7LL * leaf_base[g];
    const int n = SINGLE ? n1 : leaf_base[g + 1] - leaf_base[g];
    const int k = (int)(loc / n), i = (int)(loc % n);
    if (k == 0) {
      keep[t] = 1;
      continue;
    }
    // Element = leaf i + D[k]; it and each leaf + D[k] - D[j] (j < k) are
    // looked up in leaf i's neighbour list, except leaf + 2 D[k] (the
    // opposite direction on k's axis, j = k - 1 for even k), which goes
    // through the tree.  All lookups are issued together.
    const GridView& v = SINGLE ? sv : views[g];
    const VR_GLOBAL int* nb = v.nbr + (size_t)i * VR_NBR;
    const int kx = c_dir7[k][0] >> 3, ky = c_dir7[k][1] >> 3, kz = c_dir7[k][2] >> 3;
    int far = -1;
    if ((k & 1) == 0) {
      const int4 o = v.origins[i];
      far = find_leaf(v, o.x + 2 * c_dir7[k][0], o.y + 2 * c_dir7[k][1], o.z + 2 * c_dir7[k][2]);
    }
    int kp = nb[(kx + 1) * 9 + (ky + 1) * 3 + (kz + 1)] < 0;
    for (int j = 1; j < k; ++j) {
      const int dx = kx - (c_dir7[j][0] >> 3), dy = ky - (c_dir7[j][1] >> 3), dz = kz - (c_dir7[j][2] >> 3);
      const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
      if (near && nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] >= 0) kp = 0;
    }
    if (far >= 0) kp = 0;
    keep[t] = kp;
  }
}

// tile_base[g] = first tile of grid g (G + 1 entries; tile_base[G] = total),
// written by the first G + 1 threads of the scatter launch.
__device__ __forceinline__ void mc_tile_base(const int* leaf_base, int G, const int* pos, int nt, int* tile_base) {
  const long long total7 = 7LL * leaf_base[G];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
    const long long s = 7LL * leaf_base[g];
    tile_base[g] = s < total7 ? pos[s] : nt;
  }
}

// tanchor[rank] = i << 3 | k: the tile is leaf i of its grid plus D[k].
// Also writes tile_base.  LDS (one grid of at most MC_SCATTER_LDS leaves):
// each block first copies the origins into LDS and searches there.
#define MC_SCATTER_LDS 4096
template <bool LDS>
__global__ __launch_bounds__(256) void k_mc_scatter_tiles(const GridView* views, const int* leaf_base, int G,
                                                          const int* keep, const int* pos, const int* ntiles,
                                                          int4* tiles, int* tanchor, int* tile_base) {
  __shared__ int4 s_org[LDS ? MC_SCATTER_LDS : 1];
  const long long total = 7LL * leaf_base[G];
  const int nt = *ntiles;
  mc_tile_base(leaf_base, G, pos, nt, tile_base);
  if (LDS) {
    const int n0 = leaf_base[1] - leaf_base[0];
    for (int q = threadIdx.x; q < n0; q += blockDim.x) s_org[q] = views[0].origins[q];
    __syncthreads();
  }
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    if (!keep[t]) continue;
    const int g = segment_of(leaf_base, G, t / 7);
    const long long seg = 7LL * leaf_base[g];
    const long long loc = t - seg;
    const int n = leaf_base[g + 1] - leaf_base[g];
    const int k = (int)(loc / n), i = (int)(loc % n);
    const GridView& v = views[g];
    const int4* org = LDS ? s_org : v.origins;
    const int4 o = org[i];
    const int ex = o.x + c_dir7[k][0], ey = o.y + c_dir7[k][1], ez = o.z + c_dir7[k][2];
    // The 7 lower bounds in lockstep (the halving sequence depends on n
    // only), so each step issues 7 independent loads.
    int b[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) b[j] = 0;
    for (int len = n; len > 1;) {
      const int half = len >> 1;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int4 r = org[b[j] + half];
        if (coord_less(r.x, r.y, r.z, ex - c_dir7[j][0], ey - c_dir7[j][1], ez - c_dir7[j][2])) b[j] += half;
      }
      len -= half;
    }
    long long rank = pos[seg];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const int4 r = org[b[j]];
      const int lb = b[j] + (coord_less(r.x, r.y, r.z, ex - c_dir7[j][0], ey - c_dir7[j][1], ez - c_dir7[j][2]) ? 1 : 0);
      const long long lj = seg + (long long)j * n;
      rank += (lb < n ? pos[lj + lb] : pos[lj + n - 1] + keep[lj + n - 1]) - pos[lj];
    }
    if (rank < nt) {
      tiles[rank] = make_int4(ex, ey, ez, g);
      tanchor[rank] = (i << 3) | k;
    }
  }
}

// Per tile, over its 27 neighbour slots d = (dx + 1) * 9 + (dy + 1) * 3 +
// (dz + 1): the leaves at offsets with every component 0 or +1 (the 9^3
// corner values of the tile's cells lie in them) and the tiles at d <= 13
// (the tile itself and those before it in visiting order: every owner of
// an edge of the tile's cells lies in one; -1 when not processed).  The
// other slots are not read and hold -1.  A leaf slot within one leaf of
// the tile's anchor leaf comes from that leaf's neighbour list (all of them
// for a tile that is a leaf or lies on a leaf's -x, -y or -z side), the
// rest from the tree.
//
// A block takes MC_TN_TILES tiles, one per lane, and its four waves split
// the slots: wave w searches the tiles at d = w, w + 4, w + 8 with three
// lower bounds in lockstep over the grid's sorted tiles (neighbouring
// lanes search neighbouring origins, so the halvings share cache lines),
// and looks up two of the eight leaf slots; d = 12 (0, 0, -1) can only be
// the tile just before.  The slots are gathered in LDS and written out
// tile by tile.
// Both forms also write, per tile and cell class (27: per axis local 0 /
// 1..6 / 7), the edges a cell of that class owns itself (no candidate
// ordered before it lies in a processed tile) in bits 0..11, and in bits
// 16..27 the edges owned in an earlier tile that such a cell resolves for
// its tile (a candidate ordered before the tile's first one is processed,
// and no candidate of the tile precedes the cell).  They depend only on
// which neighbour tiles are processed, so the count and emit passes load
// them instead of deriving them per tile; candidates ordered before a cell
// lie in tiles d <= 13, the tile itself and those before it.  tt(d): the
// tile's neighbour tile at slot d (-1: not processed).
template <class TT>
__device__ __forceinline__ uint32_t mc_class_mask(int c, TT tt) {
  uint32_t processed = 0u;
#pragma unroll
  for (int d = 0; d <= 13; ++d) processed |= (uint32_t)(tt(d) >= 0) << d;
  uint32_t self = 0u, fitem = 0u;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const uint32_t b = c_mc_before[c * 12 + e], fb = c_mc_fbefore[c * 12 + e];
    self |= (uint32_t)((b & processed) == 0u) << e;
    fitem |= (uint32_t)((fb & processed) != 0u && ((b >> 13) & 1u) == 0u) << e;
  }
  return self | (fitem << 16);
}

// Slot d of tile t (k_mc_tile_nbr_slots): writes tleaf and ttile, returns the ttile entry.
__device__ __forceinline__ int mc_tile_slot(const GridView* views, const int4* tiles, const int* tanchor, const int* tile_base, int t,
                            int d, int* tleaf, int* ttile) {
  const size_t q = (size_t)t * VR_NBR + d;
  {
    const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
    const bool want_leaf = dx >= 0 && dy >= 0 && dz >= 0, want_tile = d <= 13;
    if (!want_leaf && !want_tile) {
      tleaf[q] = -1;
      ttile[q] = -1;
      return -1;
    }
    const int4 o = tiles[t];
    const int g = o.w;
    const int x = o.x + 8 * dx, y = o.y + 8 * dy, z = o.z + 8 * dz;
    int tl = -1;
    if (want_leaf) {
      const GridView& v = views[g];
      const int a = tanchor[t], k = a & 7;
      const int ax = dx + (c_dir7[k][0] >> 3), ay = dy + (c_dir7[k][1] >> 3), az = dz + (c_dir7[k][2] >> 3);
      if (ax <= 1 && ay <= 1 && az <= 1)
        tl = v.nbr[(size_t)(a >> 3) * VR_NBR + (ax + 1) * 9 + (ay + 1) * 3 + (az + 1)];
      else
        tl = find_leaf(v, x, y, z);
    }
    tleaf[q] = tl;
    int tt = -1;
    if (d == 13) {
      tt = t;
    } else if (want_tile) {
      const int lo = tile_base[g], hi = tile_base[g + 1];
      const int lb = lower_bound_origin(tiles, lo, hi, x, y, z);
      tt = (lb < hi && tiles[lb].x == x && tiles[lb].y == y && tiles[lb].z == z) ? lb : -1;
    }
    ttile[q] = tt;
    return tt;
  }
}

// One thread per slot: the lower-latency form for small tile lists (no
// search barriers; each slot's lookup runs on its own).  A block of 9 x 27
// threads takes 9 tiles at a time, so a tile's slots meet in LDS for its
// class masks.
#define MC_TNS_TILES 9
__global__ __launch_bounds__(MC_TNS_TILES * VR_NBR) void k_mc_tile_nbr_slots(
    const GridView* views, const int4* tiles, const int* tanchor, const int* ntiles, const int* tile_base, int* tleaf,
    int* ttile, uint32_t* cmask) {
  __shared__ int s_tt[MC_TNS_TILES * VR_NBR];
  const int nt = *ntiles;
  const int lt = (int)threadIdx.x / VR_NBR, d = (int)threadIdx.x % VR_NBR;
  for (int t0 = blockIdx.x * MC_TNS_TILES; t0 < nt; t0 += gridDim.x * MC_TNS_TILES) {
    const int t = t0 + lt;
    if (t < nt) s_tt[threadIdx.x] = mc_tile_slot(views, tiles, tanchor, tile_base, t, d, tleaf, ttile);
    __syncthreads();
    if (t < nt) cmask[(size_t)t * VR_NBR + d] = mc_class_mask(d, [&](int dd) { return s_tt[lt * VR_NBR + dd]; });
    __syncthreads();
  }
}

#define MC_TN_TILES 64
#define MC_TN_BIG 32768  // tile capacity from which the tile-per-lane form is used
__constant__ int c_tn_leaf[8] = {13, 14, 16, 17, 22, 23, 25, 26};
__global__ __launch_bounds__(256) void k_mc_tile_nbr(const GridView* views, const int4* tiles, const int* tanchor,
                                                     const int* ntiles, const int* tile_base, int* tleaf, int* ttile,
                                                     uint32_t* cmask) {
  __shared__ int s_tl[MC_TN_TILES * VR_NBR], s_tt[MC_TN_TILES * VR_NBR];
  const int nt = *ntiles;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t0 = blockIdx.x * MC_TN_TILES; t0 < nt; t0 += gridDim.x * MC_TN_TILES) {
    for (int i = threadIdx.x; i < MC_TN_TILES * VR_NBR; i += 256) {
      s_tl[i] = -1;
      s_tt[i] = -1;
    }
    __syncthreads();
    const int t = t0 + lane;
    if (t < nt) {
      const int4 o = tiles[t];
      const int g = o.w;
      const GridView& v = views[g];
      int* stl = s_tl + lane * VR_NBR;
      int* stt = s_tt + lane * VR_NBR;
      const int a = tanchor[t], k = a & 7;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int d = c_tn_leaf[2 * wave + j];
        const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
        const int ax = dx + (c_dir7[k][0] >> 3), ay = dy + (c_dir7[k][1] >> 3), az = dz + (c_dir7[k][2] >> 3);
        stl[d] = (ax <= 1 && ay <= 1 && az <= 1)
                     ? v.nbr[(size_t)(a >> 3) * VR_NBR + (ax + 1) * 9 + (ay + 1) * 3 + (az + 1)]
                     : find_leaf(v, o.x + 8 * dx, o.y + 8 * dy, o.z + 8 * dz);
      }
      const int lo = tile_base[g], n = tile_base[g + 1] - lo;
      int tx[3], ty[3], tz[3], b[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int d = wave + 4 * j;
        tx[j] = o.x + 8 * (d / 9 - 1);
        ty[j] = o.y + 8 * ((d / 3) % 3 - 1);
        tz[j] = o.z + 8 * (d % 3 - 1);
        b[j] = 0;
      }
      for (int len = n; len > 1;) {
        const int half = len >> 1;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int4 r = tiles[lo + b[j] + half];
          if (coord_less(r.x, r.y, r.z, tx[j], ty[j], tz[j])) b[j] += half;
        }
        len -= half;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int4 r = tiles[lo + b[j]];
        const int lb = b[j] + (coord_less(r.x, r.y, r.z, tx[j], ty[j], tz[j]) ? 1 : 0);
        bool hit = false;
        if (lb < n) {
          const int4 h = lb == b[j] ? r : tiles[lo + lb];
          hit = h.x == tx[j] && h.y == ty[j] && h.z == tz[j];
        }
        stt[wave + 4 * j] = hit ? lo + lb : -1;
      }
      if (wave == 0) {
        bool prev = false;
        if (t > lo) {
          const int4 p = tiles[t - 1];
          prev = p.x == o.x && p.y == o.y && p.z == o.z - 8;
        }
        stt[12] = prev ? t - 1 : -1;
        stt[13] = t;
      }
    }
    __syncthreads();
    const int cnt = min(MC_TN_TILES, nt - t0) * VR_NBR;
    for (int i = threadIdx.x; i < cnt; i += 256) {
      tleaf[(size_t)t0 * VR_NBR + i] = s_tl[i];
      ttile[(size_t)t0 * VR_NBR + i] = s_tt[i];
      const int* tt = s_tt + (i / VR_NBR) * VR_NBR;
      cmask[(size_t)t0 * VR_NBR + i] = mc_class_mask(i % VR_NBR, [tt](int dd) { return tt[dd]; });
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ per-cell rules
// Owner of a lattice edge: among the four cells that contain it, the one of a
// processed tile with the smallest visiting key (tile origin, lx, ly, lz).
// Within a tile the key order of the candidates of (class, edge) is fixed,
// so per tile the owner of every (class, edge) is the first candidate in
// c_mc_ownsort whose tile is processed.
__device__ __forceinline__ int mc_cls(int l) { return l == 0 ? 0 : (l == 7 ? 2 : 1); }
__device__ __forceinline__ int mc_cell_class(int x, int y, int z) { return mc_cls(x) * 9 + mc_cls(y) * 3 + mc_cls(z); }

__device__ __forceinline__ int mc_first_processed(uint32_t w, uint32_t processed) {
  int j = -1;
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    const uint32_t b = (w >> (8 * k)) & 0xFFu;
    if ((processed >> (b & 31u)) & 1u) j = (int)(b >> 5);
  }
  return j;
}

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


// ------------------------------------------------------------------ per-tile wave helpers
// Both passes run one tile per wave: a block of MC_WAVES waves works on
// MC_WAVES tiles at once, each wave on its own with its own LDS slots and no
// block barriers (a tile holds ~50 mixed cells, so a wave covers it).  Prefix
// sums in visiting order are wave scans.
#define MC_WAVES 4
#define MC_THREADS (64 * MC_WAVES)

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// First tile of worker `sub` (of nsub per block) and the stride, with the
// tiles cut into 8 contiguous ranges, one per XCD (block b runs on XCD
// b % 8); plain order when the grid has fewer than 8 blocks.
__device__ __forceinline__ int xcd_first_sub(int n, int sub, int nsub, int* end, int* stride) {
  const int B = (int)gridDim.x, b = (int)blockIdx.x;
  if (B < 8) {
    *end = n;
    *stride = B * nsub;
    return b * nsub + sub;
  }
  const int x = b & 7;
  *stride = ((B - x + 7) >> 3) * nsub;
  *end = (int)((long long)n * (x + 1) / 8);
  return (int)((long long)n * x / 8) + (b >> 3) * nsub + sub;
}

// Per-wave LDS of a tile.
struct MCTileLDS {
  uint32_t rows[81];      // sign rows: bit z of rows[x * 9 + y] = value at (x, y, z) < iso
  int tleaf[VR_NBR];      // neighbouring leaves (-1: none)
  int ttile[VR_NBR];      // neighbouring tiles (-1: not processed)
  uint16_t self[28];      // per cell class: edges the cell owns itself
  uint16_t mlist[512];    // mixed cells in visiting order
};

// Half row h (0..161) of a tile's 9^3 corner values (local 0..8): row
// r = h / 2 = x * 9 + y holds z = 0..7 from one leaf (voxels (x & 7, y & 7,
// 0..7), 32 contiguous bytes) and z = 8 from that leaf's +z neighbour; even
// h reads z = 0..3, odd h reads z = 4..7 and z = 8.  Sets rows[r] to the
// corner signs (the reference's corner test, volume_to_mesh.rs:115-123);
// the two halves sit in adjacent lanes and combine in registers, so lanes
// h and h ^ 1 must both be active.  Absent leaves read the background.
// In two steps, so that a lane's three half rows load together:
// mc_half_load issues the loads (an absent leaf's slot reads a present
// leaf or leaf 0, and the value is replaced afterwards), mc_half_bits
// combines.
struct MCHalf {
  float4 a;
  float c;
  int la, lb;
};
__device__ __forceinline__ MCHalf mc_half_load(const VR_GLOBAL NanoLeaf* leaves, const int* tl, int h) {
  const int r = h >> 1, odd = h & 1;
  const int x = (r * 57) >> 9, y = r - 9 * x;  // r / 9 for r < 81
  const int d = ((x >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + 1;
  const int row = (x & 7) * 64 + (y & 7) * 8;
  MCHalf q;
  q.la = tl[d];
  q.lb = odd ? tl[d + 1] : -1;
  q.a = *reinterpret_cast<const VR_GLOBAL float4*>(&leaves[q.la < 0 ? 0 : q.la].values[row + 4 * odd]);
  q.c = leaves[q.lb >= 0 ? q.lb : (q.la >= 0 ? q.la : 0)].values[row];
  return q;
}

// Waits here for the loads of six values at once (an empty asm that uses
// them all), so loads issued together are not each waited for at their
// first use.
template <class A, class B, class C, class D, class E, class F>
__device__ __forceinline__ void mc_use6(A a, B b, C c, D d, E e, F f) {
  asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d), "v"(e), "v"(f));
}

__device__ __forceinline__ void mc_half_bits(const MCHalf& q, float bg, float iso, int h, uint32_t* rows) {
  const int r = h >> 1, odd = h & 1;
  const float4 a = q.la >= 0 ? q.a : make_float4(bg, bg, bg, bg);
  const float c = q.lb >= 0 ? q.c : bg;
  uint32_t bits = (uint32_t)(a.x < iso) | ((uint32_t)(a.y < iso) << 1) | ((uint32_t)(a.z < iso) << 2) |
                  ((uint32_t)(a.w < iso) << 3);
  if (odd) bits = (bits << 4) | ((uint32_t)(c < iso) << 8);
  bits |= (uint32_t)__builtin_amdgcn_mov_dpp((int)bits, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]: lane ^ 1
  if (!odd) rows[r] = bits;
}

// Sign rows from the grid's sign masks (16 words per leaf), requested one
// tile ahead in three load instructions: lane l holds the leaf and +z
// neighbour words of row l, and word w1 holds, in lanes 0..16, the leaf
// word of row 64 + l and, in lanes 32..48, the +z neighbour word of row
// 64 + (l - 32) (rows < 81).  Row r = x * 9 + y has z = 0..7 in byte y & 3
// of word (x & 7) * 2 + (y & 7) / 4 of its leaf and z = 8 in the lowest bit
// of that byte in the +z neighbour; an absent leaf has the background's
// sign (mc_mask_rows).  tlr: the tile's neighbour leaves, lane d holding
// entry d (d < 27).  Whole wave.
struct MCMaskQ {
  uint32_t a0, b0, w1;  // row l's leaf and +z words; the words of rows 64.. as above
  uint32_t ok;          // bit 0: a0's leaf present, 1: b0's, 2: w1's
};
__device__ __forceinline__ MCMaskQ mc_mask_issue(const VR_GLOBAL uint32_t* sg, int tlr, int lane) {
  MCMaskQ q;
  {
    const int r = lane;
    const int x = (r * 57) >> 9, y = r - 9 * x;  // r / 9 for r < 81
    const int d = ((x >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + 1;
    const int la = __shfl(tlr, d), lb = __shfl(tlr, d + 1);
    const int wi = (x & 7) * 2 + ((y & 7) >> 2);
    q.a0 = sg[(size_t)(la < 0 ? 0 : la) * 16 + wi];
    q.b0 = sg[(size_t)(lb < 0 ? 0 : lb) * 16 + wi];
    q.ok = (uint32_t)(la >= 0) | ((uint32_t)(lb >= 0) << 1);
  }
  {
    const int hi = lane >= 32, j = (lane & 31) < 81 - 64 ? (lane & 31) : 80 - 64;
    const int r = 64 + j;
    const int x = (r * 57) >> 9, y = r - 9 * x;
    const int d = ((x >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + 1 + hi;
    const int l = __shfl(tlr, d);
    const int wi = (x & 7) * 2 + ((y & 7) >> 2);
    q.w1 = sg[(size_t)(l < 0 ? 0 : l) * 16 + wi];
    q.ok |= (uint32_t)(l >= 0) << 2;
  }
  return q;
}
// bgs: 0x1FF when the background is below iso, else 0.
__device__ __forceinline__ void mc_mask_rows(const MCMaskQ& q, uint32_t bgs, int lane, uint32_t* rows) {
  // row 64 + l's +z word and its presence from lane 32 + l
  const uint32_t b1 = (uint32_t)__shfl((int)q.w1, (lane + 32) & 63);
  const uint32_t ok1 = (uint32_t)__shfl((int)q.ok, (lane + 32) & 63);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int r = lane + 64 * k;
    const int x = (r * 57) >> 9, y = r - 9 * x;
    const int sh = (y & 3) * 8;
    const uint32_t wa = k == 0 ? q.a0 : q.w1, wb = k == 0 ? q.b0 : b1;
    const bool pa = k == 0 ? (q.ok & 1u) != 0u : (q.ok & 4u) != 0u;
    const bool pb = k == 0 ? (q.ok & 2u) != 0u : (ok1 & 4u) != 0u;
    const uint32_t lo = pa ? (wa >> sh) & 0xFFu : bgs & 0xFFu;
    const uint32_t hi = pb ? (wb >> sh) & 1u : bgs >> 8;
    if (k == 0 || lane < 81 - 64) rows[r] = lo | (hi << 8);
  }
}

// Loads the tile's neighbour lists and class masks (mc_class_mask;
// fetched into lanes < 27) into LDS, stages its sign rows (MASKS: from the
// mask words q, else from the values).  Returns the lane's mixed cells
// (corners on both sides of iso: a case other than 0 or 255): bit z of lane
// x * 8 + y is cell (x, y, z).  W: per-wave LDS with rows, tleaf, ttile and
// self (and fitem when FITEM).  Whole wave, uniform control flow.
template <bool MASKS, bool FITEM = false, class W>
__device__ __forceinline__ uint32_t mc_tile_rows(const VR_GLOBAL NanoLeaf* leaves, const MCMaskQ& q, float bg,
                                                 float iso, int tl, int tt, uint32_t cm, W& w) {
  const int lane = lane_id();
  if (lane < VR_NBR) {
    w.tleaf[lane] = tl;
    w.ttile[lane] = tt;
    w.self[lane] = (uint16_t)(cm & 0xFFFFu);
    if constexpr (FITEM) w.fitem[lane] = (uint16_t)(cm >> 16);
  }
  wave_lds_sync();
  if constexpr (MASKS) {
    mc_mask_rows(q, bg < iso ? 0x1FFu : 0u, lane, w.rows);
  } else {
    const MCHalf q0 = mc_half_load(leaves, w.tleaf, lane), q1 = mc_half_load(leaves, w.tleaf, lane + 64);
    const MCHalf q2 = mc_half_load(leaves, w.tleaf, lane < 162 - 128 ? lane + 128 : 161);
    mc_use6(q0.a.w, q1.a.w, q2.a.w, q0.c, q1.c, q2.c);
    mc_half_bits(q0, bg, iso, lane, w.rows);
    mc_half_bits(q1, bg, iso, lane + 64, w.rows);
    if (lane < 162 - 128) mc_half_bits(q2, bg, iso, lane + 128, w.rows);
  }
  wave_lds_sync();
  // Lane x * 8 + y tests its 8 cells along z with bit operations on the
  // four sign rows around it.
  const int x = lane >> 3, y = lane & 7;
  const uint32_t r00 = w.rows[x * 9 + y], r10 = w.rows[(x + 1) * 9 + y];
  const uint32_t r01 = w.rows[x * 9 + y + 1], r11 = w.rows[(x + 1) * 9 + y + 1];
  const uint32_t all = r00 & r10 & r01 & r11, any = r00 | r10 | r01 | r11;
  const uint32_t one = all & (all >> 1), zero = ~any & ~(any >> 1);
  return ~(one | zero) & 0xFFu;
}

// mc_tile_rows, then the tile's mixed cells in visiting order in w.mlist.
// Returns the mixed-cell count.
template <bool MASKS>
__device__ __forceinline__ int mc_tile_setup(const VR_GLOBAL NanoLeaf* leaves, const MCMaskQ& q, float bg, float iso,
                                             int tl, int tt, uint32_t cm, MCTileLDS& w) {
  const int lane = lane_id();
  uint32_t mixed = mc_tile_rows<MASKS>(leaves, q, bg, iso, tl, tt, cm, w);
  const int cnt = __popc(mixed);
  const int incl = wave_incl_scan(cnt);
  int k = incl - cnt;
  for (; mixed; mixed &= mixed - 1u, ++k) w.mlist[k] = (uint16_t)(lane * 8 + (__ffs(mixed) - 1));
  wave_lds_sync();
  return __builtin_amdgcn_readlane(incl, 63);
}

// Cube index of cell c = x * 64 + y * 8 + z from the sign rows, corners in
// the reference's order (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1)
// (1,1,1) (0,1,1).
__device__ __forceinline__ uint32_t mc_case_rows(const uint32_t* rows, int x, int y, int z) {
  const uint32_t r00 = rows[x * 9 + y] >> z, r10 = rows[(x + 1) * 9 + y] >> z;
  const uint32_t r11 = rows[(x + 1) * 9 + y + 1] >> z, r01 = rows[x * 9 + y + 1] >> z;
  return (r00 & 1u) | ((r10 & 1u) << 1) | ((r11 & 1u) << 2) | ((r01 & 1u) << 3) | (((r00 >> 1) & 1u) << 4) |
         (((r10 >> 1) & 1u) << 5) | (((r11 >> 1) & 1u) << 6) | (((r01 >> 1) & 1u) << 7);
}

// Byte offset of local edge e of cell (x, y, z) in the tile's lattice-edge
// table of 16-bit entries (es = c_mc_eslot[e]: twice the slot base | twice
// the x stride << 12 | twice the y stride << 20, so no shift remains):
// every cell containing the same lattice edge gets the same slot.
__device__ __forceinline__ uint32_t mc_edge_off(int x, int y, int z, uint32_t es) {
  return (es & 0xFFFu) + (uint32_t)x * ((es >> 12) & 0xFFu) + (uint32_t)y * (es >> 20) + 2u * (uint32_t)z;
}
__device__ __forceinline__ uint16_t& mc_edge_at(uint16_t* edge, uint32_t off) {
  return *reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(edge) + off);
}
__device__ __forceinline__ uint32_t mc_edge_get(const uint16_t* edge, uint32_t off) {
  return *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(edge) + off);
}
#define MC_EDGES 1944

// Index (relative to the grid) of the vertex on local edge e of cell
// (x, y, z) whose owner lies in an earlier tile: the first processed
// candidate in visiting order (c_mc_ownsort) names that tile and cell, whose
// face meta from pass 1 gives the vertex.  Such an owner always has a face
// index: its tile precedes the cell's, so its first differing tile offset is
// -1 and its local coordinate on that axis is 7.
__device__ __forceinline__ int mc_foreign_index(int x, int y, int z, int e, uint32_t processed,
                                                const uint32_t* s_ownsort, const uint32_t* s_cand, const int* ttile,
                                                const uint32_t* meta, const int* vbase, int vrel) {
  const int j = mc_first_processed(s_ownsort[mc_cell_class(x, y, z) * 12 + e], processed);
  const uint32_t cw = s_cand[e * 4 + (j < 0 ? 0 : j)];
  const int nx = x + (int)(int8_t)(cw & 0xFFu), ny = y + (int)(int8_t)((cw >> 8) & 0xFFu),
            nz = z + (int)(int8_t)((cw >> 16) & 0xFFu);
  const uint32_t ep = (uint32_t)(int8_t)(cw >> 24);
  const int d = ((nx >> 3) + 1) * 9 + ((ny >> 3) + 1) * 3 + ((nz >> 3) + 1);
  const int ot = ttile[d];
  const int fi = mc_face_index(nx & 7, ny & 7, nz & 7);
  const bool ok = j >= 0 && d != 13 && ot >= 0 && fi >= 0;
  const uint32_t om = meta[(size_t)(ok ? ot : 0) * MC_FACE + (ok ? fi : 0)];
  const int ob = vbase[ok ? ot : 0];
  return ok ? ob - vrel + (int)(om >> 12) + __popc(om & 0xFFFu & ((1u << ep) - 1u)) : 0;
}

// The two corner values of local edge e of cell (x, y, z), requested with
// their leaves' presence (an absent leaf's slot reads leaf 0 and is
// replaced by the background in mc_vertex_store), so that a round's loads
// can be issued together and consumed later.
struct MCCorner {
  float va, vb;
  int ok;  // bit 0: corner a's leaf present, bit 1: corner b's
};
// A tile lattice point (x, y, z in 0..8) as packed fields x << 8 | y << 4 | z.
__device__ __forceinline__ uint32_t mc_pack(int x, int y, int z) { return ((uint32_t)x << 8) | ((uint32_t)y << 4) | (uint32_t)z; }
// Its leaf: the tile's own (slot 13) or the +x / +y / +z neighbour once that
// coordinate is 8 (bit 3 of the field).
__device__ __forceinline__ int mc_slot9(uint32_t s) {
  return 13 + 9 * (int)((s >> 11) & 1u) + 3 * (int)((s >> 7) & 1u) + (int)((s >> 3) & 1u);
}
// Its offset in that leaf (x & 7) * 64 + (y & 7) * 8 + (z & 7).
__device__ __forceinline__ int mc_off9(uint32_t s) { return (int)(((s >> 2) & 0x1C0u) | ((s >> 1) & 0x38u) | (s & 7u)); }

__device__ __forceinline__ MCCorner mc_corner_load(const VR_GLOBAL NanoLeaf* leaves, const int* tl, int x, int y, int z,
                                                   int e, const uint32_t* s_einfo) {
  const uint32_t ei = s_einfo[e], p = mc_pack(x, y, z);
  const uint32_t sa = p + (ei & 0xFFFFu), sb = p + (ei >> 16);
  const int la = tl[mc_slot9(sa)], lb = tl[mc_slot9(sb)];
  MCCorner c;
  c.va = leaves[la < 0 ? 0 : la].values[mc_off9(sa)];
  c.vb = leaves[lb < 0 ? 0 : lb].values[mc_off9(sb)];
  c.ok = (int)(la >= 0) | ((int)(lb >= 0) << 1);
  return c;
}

// Vertex on owned local edge e of cell (x, y, z) (volume_to_mesh.rs:162-186)
// from its corner values: t = (iso - va) / (vb - va) in f32 (0.5 when
// |vb - va| <= 1e-10), position pa + t (pb - pa) in f64 from the corners'
// world positions (index * vs).  An edge spans one axis; on the other two
// pb - pa is +0 exactly, so pa + t (pb - pa) is pa there and only the
// spanned axis is interpolated.
__device__ __forceinline__ void mc_vertex_store(const MCCorner& cv, float bg, int x, int y, int z, int e,
                                                const int4& to, double vs, float iso, const uint32_t* s_einfo,
                                                double* out) {
  const uint32_t ei = s_einfo[e], pk = mc_pack(x, y, z);
  const uint32_t sa = pk + (ei & 0xFFFFu), sb = pk + (ei >> 16);
  const int ax = (int)((sa >> 8) & 15u), ay = (int)((sa >> 4) & 15u), az = (int)(sa & 15u);
  const float va = (cv.ok & 1) ? cv.va : bg;
  const float vbv = (cv.ok & 2) ? cv.vb : bg;
  double tpar;
  if (fabsf(vbv - va) > 1e-10f)
    tpar = (double)((iso - va) / (vbv - va));
  else
    tpar = 0.5;
  const uint32_t span = (ei & 0xFFFFu) ^ (ei >> 16);  // 0x100, 0x10 or 1: the axis of the edge
  const double pax = (double)(to.x + ax) * vs, pay = (double)(to.y + ay) * vs, paz = (double)(to.z + az) * vs;
  // pa on the spanned axis is that axis's corner-a coordinate
  const bool sx = span == 0x100u, sy = span == 0x10u, sz = span == 1u;
  const int ib = sx ? to.x + (int)((sb >> 8) & 15u) : (sy ? to.y + (int)((sb >> 4) & 15u) : to.z + (int)(sb & 15u));
  const double pa = sx ? pax : (sy ? pay : paz), pb = (double)ib * vs;
  const double p = pa + tpar * (pb - pa);
  out[0] = sx ? p : pax;
  out[1] = sy ? p : pay;
  out[2] = sz ? p : paz;
}

__device__ __forceinline__ void mc_write_vertex(const VR_GLOBAL NanoLeaf* leaves, float bg, const int* tl, int x, int y,
                                                int z, int e, const int4& to, double vs, float iso,
                                                const uint32_t* s_einfo, double* out) {
  mc_vertex_store(mc_corner_load(leaves, tl, x, y, z, e, s_einfo), bg, x, y, z, e, to, vs, iso, s_einfo, out);
}

// ------------------------------------------------------------------ pass 1
// Tile t's lists in one vector load per array: lanes < 27 get its neighbour
// leaves (*l) and class masks (*k), lane 27 of *l its grid (tiles[t].w).
// (The count pass needs no neighbour tiles: the class masks carry what it
// uses of them.)
__device__ __forceinline__ void mc_count_lists(const int* tleaf, const uint32_t* cmask, const int4* tiles, int t,
                                               int lane, int* l, uint32_t* k) {
  const int* p = lane < VR_NBR ? tleaf + (size_t)t * VR_NBR + lane : reinterpret_cast<const int*>(tiles) + (size_t)t * 4 + 3;
  if (lane <= VR_NBR) *l = p[0];
  if (lane < VR_NBR) *k = cmask[(size_t)t * VR_NBR + lane];
}

// Per mixed cell (in visiting order): case, owned crossing edges, triangle
// count.  The face cells (mc_face_index >= 0) store meta = (vertex prefix
// inside the tile << 12) | owned-edge mask for the emit pass of later tiles;
// per-tile vertex and triangle totals.  Cells with case 0 or 255 add nothing
// and are never visited.  Each wave loops over tiles; the next tile's
// neighbour lists load while the current one is processed.
// LIST: the tile's mixed cells in visiting order also go to mcell as cell |
// case << 9.  Each wave writes its tiles' lists one after another into its
// own equal share of mcap cells (the tiles of a wave are fixed by the grid
// size), so no offset is shared; *mctr (64-bit, cleared before the launch)
// ends at the largest share any wave needed.  mtile[2 t] = the list's
// offset (-1 past the wave's share, and for tiles without mixed cells),
// mtile[2 t + 1] = the mixed-cell count.
template <bool MASKS, bool LIST>
__global__ __launch_bounds__(MC_THREADS) void k_mc_count(const GridView* views, const int4* tiles, const int* ntiles,
                                                         const int* tleaf, const uint32_t* cmask,
                                                         float iso, uint32_t* meta, int* tv, int* tt, int tcap,
                                                         uint32_t* mcell, int* mtile, unsigned long long* mctr,
                                                         long long mcap) {
  __shared__ uint16_t s_emn[256];
  __shared__ uint8_t s_ccls[512];  // class of cell c = x * 64 + y * 8 + z (mc_cell_class)
  __shared__ MCTileLDS s_w[MC_WAVES];
  const int tid = threadIdx.x, lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 256; i += MC_THREADS) s_emn[i] = c_mc_emn[i];
  for (int i = tid; i < 512; i += MC_THREADS) s_ccls[i] = (uint8_t)mc_cell_class(i >> 6, (i >> 3) & 7, i & 7);
  __syncthreads();
  MCTileLDS& w = s_w[wv];
  const int nt = *ntiles;
  // Totals past the tile count (tcap is a host upper bound) are cleared for the scans.
  for (int i = nt + (int)(blockIdx.x * MC_THREADS) + tid; i < tcap; i += gridDim.x * MC_THREADS) {
    tv[i] = 0;
    tt[i] = 0;
  }
  int t_end, t_step;
  int t = xcd_first_sub(nt, wv, MC_WAVES, &t_end, &t_step);
  // LIST: this wave's share of mcell and the cells its tiles used so far
  const long long share = LIST ? mcap / ((long long)gridDim.x * MC_WAVES) : 0;
  const long long share0 = share * ((long long)blockIdx.x * MC_WAVES + wv);
  long long used = 0;
  // Lists of tile t (ntl, ntt) and of the tile after it (ftl, ftt) are in
  // flight, loaded by mc_count_lists; MASKS: so are tile t's sign-mask words
  // (q).  Vector loads throughout: a scalar load of a later tile would hold
  // up every LDS wait of this one.
  int ntl = -1, ftl = -1;
  uint32_t ncm = 0u, fcm = 0u;
  if (t < t_end) mc_count_lists(tleaf, cmask, tiles, t, lane, &ntl, &ncm);
  if (t + t_step < t_end) mc_count_lists(tleaf, cmask, tiles, t + t_step, lane, &ftl, &fcm);
  MCMaskQ q{};
  if (MASKS && t < t_end) q = mc_mask_issue(views[__builtin_amdgcn_readlane(ntl, VR_NBR)].signs, ntl, lane);
  for (; t < t_end; t += t_step) {
    const int tl = ntl;
    const uint32_t cm = ncm;
    const int tn = t + t_step, tf = tn + t_step;
    ntl = ftl;
    ncm = fcm;
    if (tf < t_end) mc_count_lists(tleaf, cmask, tiles, tf, lane, &ftl, &fcm);
    const GridView& g = views[__builtin_amdgcn_readlane(tl, VR_NBR)];
    const MCMaskQ cq = q;
    const int n = mc_tile_setup<MASKS>(g.leaves, cq, g.background, iso, tl, -1, cm, w);
    // the next tile's mask words load while this one is counted
    if (MASKS && tn < t_end) q = mc_mask_issue(views[__builtin_amdgcn_readlane(ntl, VR_NBR)].signs, ntl, lane);
    long long moff = -1;
    if (LIST && n > 0) {
      moff = used + n <= share ? share0 + used : -1;
      used += n;
    }
    int vcarry = 0, tcarry = 0;
    for (int base = 0; base < n; base += 64) {
      const int i = base + lane;
      int c = 0, packed = 0;
      uint32_t owned = 0u;
      if (i < n) {
        c = w.mlist[i];
        const int x = c >> 6, y = (c >> 3) & 7, z = c & 7;
        const uint32_t cube = mc_case_rows(w.rows, x, y, z);
        const uint32_t emn = s_emn[cube];
        owned = (emn & 0xFFFu) & w.self[s_ccls[c]];
        packed = __popc(owned) | (int)((emn >> 12) << 16);
        if (LIST && moff >= 0) mcell[moff + i] = (uint32_t)c | (cube << 9);
      }
      const int incl = wave_incl_scan(packed);
      if (i < n) {
        const int fi = mc_face_index(c >> 6, (c >> 3) & 7, c & 7);
        if (fi >= 0) meta[(size_t)t * MC_FACE + fi] = ((uint32_t)(vcarry + ((incl - packed) & 0xFFFF)) << 12) | owned;
      }
      const int tot = __builtin_amdgcn_readlane(incl, 63);
      vcarry += tot & 0xFFFF;
      tcarry += tot >> 16;
    }
    // the tile's totals (and LIST: its list offset and count) in one store
    // instruction, one lane per array
    if (lane < (LIST ? 4 : 2)) {
      int* const dst = lane == 0 ? tv + t : (lane == 1 ? tt + t : mtile + 2 * (size_t)t + (lane - 2));
      *dst = lane == 0 ? vcarry : (lane == 1 ? tcarry : (lane == 2 ? (int)moff : n));
    }
    wave_lds_sync();
  }
  if (LIST && lane == 0 && used > 0) atomicMax(mctr, (unsigned long long)used);
}

// Per-grid vertex and triangle ranges: out[4 g] = first vertex, [4 g + 1] =
// vertex count, [4 g + 2] = first triangle, [4 g + 3] = triangle count.
// counts_out (optional): the vertex and triangle totals.  Written by the
// first block of the emit launch (the scans are complete by then).
__device__ __forceinline__ void mc_ranges(const int* tile_base, int G, int nt, const int* vbase, const int* tbase,
                                          const int* totals, int* out, int* counts_out) {
  if (counts_out && threadIdx.x == 0) {
    counts_out[0] = totals[1];
    counts_out[1] = totals[2];
  }
  for (int g = threadIdx.x; g < G; g += blockDim.x) {
    const int a = tile_base[g], b = tile_base[g + 1];
    const int va = a < nt ? vbase[a] : totals[1], vb = b < nt ? vbase[b] : totals[1];
    const int ta = a < nt ? tbase[a] : totals[2], tb = b < nt ? tbase[b] : totals[2];
    out[4 * g] = va;
    out[4 * g + 1] = vb - va;
    out[4 * g + 2] = ta;
    out[4 * g + 3] = tb - ta;
  }
}

__global__ void k_mc_ranges(const int* tile_base, int G, const int* ntiles, const int* vbase, const int* tbase,
                            const int* totals, int* out, int* counts_out) {
  mc_ranges(tile_base, G, *ntiles, vbase, tbase, totals, out, counts_out);
}

// ------------------------------------------------------------------ pass 2
// Owned vertices (owner's edge orientation, f32 t, f64 position;
// volume_to_mesh.rs:162-186) and triangles in TRI_TABLE order (:191-200).
// Mixed cells go in rounds of 64, one per lane (with the count pass's
// lists: of 128, two consecutive cells per lane): a wave scan places the
// round's vertices and triangles, and lists of the round's vertices
// (cell, edge) and triangles (round cell and row entry; with the count
// pass's lists, the cell and the triangle's three edges) spread the writes
// one item per lane.  Vertex indices go through the tile's lattice-edge table
// in LDS (16 bits per edge): an owned vertex stores its index relative to
// the tile's first vertex at its edge's slot; an edge owned in an earlier
// tile (it lies on one of the tile's three low planes, at most 3 x 144) is
// resolved once, by the first cell of this tile that contains it, from that
// tile's face meta of pass 1: into a side table of MC_FSIDE entries, its
// slot holding 0x8000 | the entry, or, past the side table, its slot holding
// 0xFFFF and every triangle corner on it resolving it again.  Owners
// precede the cells that share their edges in visiting order, so every
// slot a triangle reads was written in this or an earlier round, before the
// round's triangles.  A round with more items than the lists hold is
// written cell by cell.  Vertex indices are relative to the tile's grid.
// The per-wave LDS is kept to 6.1 KB (the round's cell list is built per
// round, the 32-bit triangle entries of the list mode share the space of
// the sign rows and cell lists it does not use, the triangle table is
// nibble-packed) and a block's to 53 744 bytes, so that three 8-wave
// blocks fit a CU's 160 KB (6 waves per SIMD): LDS is granted in units
// coarser than the byte count, and at 54 256 bytes per block only two
// blocks were resident (the emit pass took 10.8 ms instead of 8.0).
#define MC_EWAVES 8
#define MC_ETHREADS (64 * MC_EWAVES)
#define MC_RV 128   // vertices listed per round (rounds of 64 cells)
#define MC_RT 192   // triangles listed per round (rounds of 64 cells)
#define MC_RVL 192  // vertices listed per round of 128 cells (LIST)
#define MC_RTL 288  // triangles listed per round of 128 cells (LIST)
#define MC_FL 192   // earlier-tile edges listed per resolution pass
#define MC_FSIDE 96 // side-table entries per tile
struct MCEmitLDS {
  union {
    struct {                     // classification from sign rows (no lists)
      uint32_t rows[81];         // sign rows: bit z of rows[x * 9 + y] = value at (x, y, z) < iso
      union {
        uint16_t tlist[MC_RT];   // round cell | triangle of the cell << 6
        uint16_t flist[MC_FL];   // cell << 4 | edge, earlier-tile edges to resolve
      };
      uint16_t rlist[64];        // the round's mixed cells
      uint8_t rcube[64];         // their cases
    };
    union {                      // lists from the count pass
      uint32_t tlist32[MC_RTL];  // cell | the triangle's edges << 9 (4 bits each)
      uint16_t flist32[MC_FL];   // as flist
    };
  };
  int tleaf[VR_NBR];             // neighbouring leaves (-1: none)
  int ttile[VR_NBR];             // neighbouring tiles (-1: not processed)
  uint16_t self[28];             // per cell class: edges the cell owns itself
  uint16_t fitem[28];            // per cell class: edges owned in an earlier tile that this cell resolves
  uint16_t edge[MC_EDGES];       // lattice-edge table of the tile
  union {
    struct {
      int foreign[MC_FSIDE];     // vertex index (relative to the grid) of edges owned in an earlier tile
      uint16_t vlist[MC_RVL];    // cell << 4 | edge
    };
    // the side table read at any 7-bit slot (an entry's low 7 bits), so a
    // lookup needs no range check; slots past MC_FSIDE read the vertex list,
    // and their values are never used
    int fwin[128];
  };
};
static_assert(MC_FSIDE <= 128 && MC_FSIDE * 4 + MC_RVL * 2 >= 128 * 4, "side-table window inside the LDS of a wave");

// Tile t's lists and header in one vector load per array: lanes < 27 get
// its neighbour leaves (*l), tiles (*m) and class masks (*k); lanes 27..34
// of *l its vertex and triangle counts, first vertex and triangle, and
// tiles[t] (x, y, z, grid); LIST: lanes 35 and 36 its mixed-cell list
// offset and count (mtile).
template <bool LIST>
__device__ __forceinline__ void mc_emit_lists(const int* tleaf, const int* ttile, const uint32_t* cmask, const int* tv,
                                              const int* tt, const int* vbase, const int* tbase, const int4* tiles,
                                              const int* mtile, int t, int lane, int* l, int* m, uint32_t* k) {
  const int* p = lane < VR_NBR ? tleaf + (size_t)t * VR_NBR + lane
                 : lane == VR_NBR     ? tv + t
                 : lane == VR_NBR + 1 ? tt + t
                 : lane == VR_NBR + 2 ? vbase + t
                 : lane == VR_NBR + 3 ? tbase + t
                 : lane < VR_NBR + 8  ? reinterpret_cast<const int*>(tiles) + (size_t)t * 4 + (lane - VR_NBR - 4)
                                      : mtile + 2 * (size_t)t + (lane - VR_NBR - 8);
  if (lane < VR_NBR + (LIST ? 10 : 8)) *l = p[0];
  if (lane < VR_NBR) {
    *m = ttile[(size_t)t * VR_NBR + lane];
    *k = cmask[(size_t)t * VR_NBR + lane];
  }
}

// The round's cells (visiting positions base .. base + 63) into rlist.
// *rem: the lane's mixed cells not yet listed, *rk: the visiting position
// of the first of them; rounds run in order, so a lane lists only its cells
// of this round and consumes them.
__device__ __forceinline__ void mc_round_list(uint32_t* rem, int* rk, int base, int lane, uint16_t* rlist) {
  uint32_t m = *rem;
  int k = *rk;
  for (; m && k < base + 64; m &= m - 1u, ++k) rlist[k - base] = (uint16_t)(lane * 8 + (__ffs(m) - 1));
  *rem = m;
  *rk = k;
}

// LIST: requests the entries of the first round (128 cells) of the tile
// whose header hl holds (mc_emit_lists): *a = entry 2 lane, *b = entry
// 2 lane + 1 (0 past the count).
__device__ __forceinline__ void mc_list_issue(const uint32_t* mcell, int hl, int lane, uint32_t* a, uint32_t* b) {
  const int off = __builtin_amdgcn_readlane(hl, VR_NBR + 8), cnt = __builtin_amdgcn_readlane(hl, VR_NBR + 9);
  const int nv = __builtin_amdgcn_readlane(hl, VR_NBR), ntr = __builtin_amdgcn_readlane(hl, VR_NBR + 1);
  if (off < 0 || (nv == 0 && ntr == 0)) return;
  if (2 * lane < cnt) *a = mcell[(size_t)off + 2 * lane];
  if (2 * lane + 1 < cnt) *b = mcell[(size_t)off + 2 * lane + 1];
}

// What resolving an earlier tile's edge again needs (mc_foreign_index).
struct MCForeign {
  uint32_t processed;
  const uint32_t* s_ownsort;
  const uint32_t* s_cand;
  const int* ttile;
  const uint32_t* meta;
  const int* vbase;
  int vrel;
};

// Vertex index (relative to the grid) from an edge-table entry v other than
// 0xFFFF: 0x8000 | side-table slot, or the vertex relative to the tile's
// first (vbrel = that vertex relative to the grid's first).
__device__ __forceinline__ int mc_corner_known(uint32_t v, const MCEmitLDS& el, int vbrel) {
  const int f = el.fwin[v & 0x7Fu];
  return (v & 0x8000u) ? f : vbrel + (int)v;
}

// Vertex index (relative to the grid) on local edge e of cell (x, y, z).
__device__ __forceinline__ int mc_corner_index(int x, int y, int z, int e, const MCEmitLDS& el,
                                               const uint32_t* s_eslot, int vbrel, const MCForeign& fc) {
  const uint32_t v = mc_edge_get(el.edge, mc_edge_off(x, y, z, s_eslot[e]));
  if (v == 0xFFFFu)
    return mc_foreign_index(x, y, z, e, fc.processed, fc.s_ownsort, fc.s_cand, fc.ttile, fc.meta, fc.vbase, fc.vrel);
  const int f = el.fwin[v & 0x7Fu];
  return (v & 0x8000u) ? f : vbrel + (int)v;
}

// LIST: the tiles' mixed cells and cases come from the count pass's lists
// (mcell, mtile) instead of the sign rows; the first round's entries of
// the next tile load while the current one is processed.  A tile whose
// list did not fit (offset -1) writes nothing: that happens only when the
// mesh is larger than the capacity the lists were sized for.
// Edges owned in an earlier tile (fm0 of cell c0, then fm1 of cell c1, per
// lane), one resolved per lane, MC_FL per pass: into the side table from
// slot *fcarry on (its edge slot holding 0x8000 | the entry), or past the
// side table marked 0xFFFF and resolved again where a triangle needs it.
// Each lattice edge is listed once per tile (by the first cell of the tile
// containing it).  Whole wave.
__device__ __forceinline__ void mc_foreign_pass(uint32_t fm0, int c0, uint32_t fm1, int c1, int lane, uint16_t* flist,
                                                int* fcarry, MCEmitLDS& el, const uint32_t* s_eslot,
                                                const MCForeign& fc) {
  const int nf = __popc(fm0) + __popc(fm1);
  const int fincl = wave_incl_scan(nf);
  const int ftot = __builtin_amdgcn_readlane(fincl, 63);
  for (int fb = 0; fb < ftot; fb += MC_FL) {
    int k = fincl - nf;
    for (uint32_t m = fm0; m; m &= m - 1u, ++k)
      if (k >= fb && k < fb + MC_FL) flist[k - fb] = (uint16_t)((c0 << 4) | (__ffs(m) - 1));
    for (uint32_t m = fm1; m; m &= m - 1u, ++k)
      if (k >= fb && k < fb + MC_FL) flist[k - fb] = (uint16_t)((c1 << 4) | (__ffs(m) - 1));
    wave_lds_sync();
    const int fn = ftot - fb < MC_FL ? ftot - fb : MC_FL;
    for (int k2 = lane; k2 < fn; k2 += 64) {
      const int it = flist[k2], cc = it >> 4, fe = it & 15;
      const int fx = cc >> 6, fy = (cc >> 3) & 7, fz = cc & 7;
      const int slot = *fcarry + fb + k2;
      const uint32_t es = mc_edge_off(fx, fy, fz, s_eslot[fe]);
      if (slot < MC_FSIDE) {
        el.foreign[slot] =
            mc_foreign_index(fx, fy, fz, fe, fc.processed, fc.s_ownsort, fc.s_cand, fc.ttile, fc.meta, fc.vbase, fc.vrel);
        mc_edge_at(el.edge, es) = (uint16_t)(0x8000u | (uint32_t)slot);
      } else {
        mc_edge_at(el.edge, es) = 0xFFFFu;
      }
    }
    wave_lds_sync();
  }
  *fcarry += ftot;
}

template <bool MASKS, bool LIST>
__global__ __launch_bounds__(MC_ETHREADS) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_mc_emit(
    const GridView* views, const int4* tiles, const int* ntiles, const int* tleaf, const int* ttile,
    const int* tile_base, const uint32_t* cmask, float iso, const uint32_t* meta, const int* tv, const int* tt,
    const int* vbase, const int* tbase, const int* totals, double* verts, uint32_t* tris, long long vcap,
    long long tricap, int G, int* ranges, int* counts_out, const uint32_t* mcell, const int* mtile) {
  __shared__ uint16_t s_emn[256];
  __shared__ uint64_t s_trin[256];  // TRI_TABLE rows, one nibble per entry
  __shared__ uint32_t s_ownsort[27 * 12];
  __shared__ uint8_t s_ccls[512];  // class of cell c = x * 64 + y * 8 + z (mc_cell_class)
  __shared__ uint32_t s_cand[48];
  __shared__ uint32_t s_einfo[12];
  __shared__ uint32_t s_eslot[12];
  __shared__ MCEmitLDS s_e[MC_EWAVES];
  const int tid = threadIdx.x, lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 256; i += MC_ETHREADS) {
    s_emn[i] = c_mc_emn[i];
    s_trin[i] = c_mc_trin[i];
  }
  for (int i = tid; i < 27 * 12; i += MC_ETHREADS) s_ownsort[i] = c_mc_ownsort[i];
  for (int i = tid; i < 512; i += MC_ETHREADS) s_ccls[i] = (uint8_t)mc_cell_class(i >> 6, (i >> 3) & 7, i & 7);
  if (tid < 48) s_cand[tid] = c_mc_cand4[tid];
  if (tid < 12) {
    s_einfo[tid] = c_mc_edgeinfo[tid];
    s_eslot[tid] = c_mc_eslot[tid];
  }
  __syncthreads();
  MCEmitLDS& el = s_e[wv];
  const int nt = *ntiles;
  // LIST: the ranges come from k_mc_ranges before this launch, so a tile's
  // first grid vertex is a scalar-cache read of ranges[4 g]
  if (!LIST && blockIdx.x == 0) mc_ranges(tile_base, G, nt, vbase, tbase, totals, ranges, counts_out);
  // writes stop at the counts or at the output capacity, whichever is lower
  const int vmax = (int)min((long long)totals[1], vcap), tmax = (int)min((long long)totals[2], tricap);
  int t_end, t_step;
  int t = xcd_first_sub(nt, wv, MC_EWAVES, &t_end, &t_step);
  // Lists and header of tile t (h_tl, h_tt) and of the tile after it (ftl,
  // ftt) are in flight, loaded by mc_emit_lists, so that the next tile's
  // sign-mask words (MASKS, q) can be requested in the middle of the
  // current tile.  Vector loads throughout: a scalar load of a later tile
  // would hold up every LDS wait of this one.
  int h_tl = -1, h_tt = -1, ftl = -1, ftt = -1;
  uint32_t h_cm = 0u, fcm = 0u;
  if (t < t_end)
    mc_emit_lists<LIST>(tleaf, ttile, cmask, tv, tt, vbase, tbase, tiles, mtile, t, lane, &h_tl, &h_tt, &h_cm);
  if (t + t_step < t_end)
    mc_emit_lists<LIST>(tleaf, ttile, cmask, tv, tt, vbase, tbase, tiles, mtile, t + t_step, lane, &ftl, &ftt, &fcm);
  MCMaskQ q{};
  if (MASKS && !LIST && t < t_end)
    q = mc_mask_issue(views[__builtin_amdgcn_readlane(h_tl, VR_NBR + 7)].signs, h_tl, lane);
  // LIST: entries 2 lane and 2 lane + 1 of tile t (e0, e1: its first
  // round); lanes past the count hold 0.
  uint32_t e0 = 0u, e1 = 0u;
  if (LIST && t < t_end) mc_list_issue(mcell, h_tl, lane, &e0, &e1);
  for (; t < t_end; t += t_step) {
    const int tl = h_tl, ttv = h_tt;
    const uint32_t cm = h_cm;
    const int c_nv = __builtin_amdgcn_readlane(tl, VR_NBR), c_nt = __builtin_amdgcn_readlane(tl, VR_NBR + 1);
    const int vb = __builtin_amdgcn_readlane(tl, VR_NBR + 2), tb = __builtin_amdgcn_readlane(tl, VR_NBR + 3);
    const int4 to = make_int4(__builtin_amdgcn_readlane(tl, VR_NBR + 4), __builtin_amdgcn_readlane(tl, VR_NBR + 5),
                              __builtin_amdgcn_readlane(tl, VR_NBR + 6), __builtin_amdgcn_readlane(tl, VR_NBR + 7));
    const int tn = t + t_step, tf = tn + t_step;
    const int moff = LIST ? __builtin_amdgcn_readlane(tl, VR_NBR + 8) : 0;
    const int mn = LIST ? __builtin_amdgcn_readlane(tl, VR_NBR + 9) : 0;
    h_tl = ftl;
    h_tt = ftt;
    h_cm = fcm;
    if (tf < t_end)
      mc_emit_lists<LIST>(tleaf, ttile, cmask, tv, tt, vbase, tbase, tiles, mtile, tf, lane, &ftl, &ftt, &fcm);
    const MCMaskQ cq = q;
    uint32_t c0 = e0, c1 = e1;  // LIST: this tile's first round
    if (LIST) {
      e0 = e1 = 0u;
      if (tn < t_end) mc_list_issue(mcell, h_tl, lane, &e0, &e1);
    }
    if (c_nv == 0 && c_nt == 0) {  // nothing to write (uniform over the wave)
      if (MASKS && !LIST && tn < t_end)
        q = mc_mask_issue(views[__builtin_amdgcn_readlane(h_tl, VR_NBR + 7)].signs, h_tl, lane);
      continue;
    }
    if (LIST && moff < 0) continue;  // the list did not fit: the mesh exceeds its capacity
    const uint32_t processed = (uint32_t)__ballot(lane < VR_NBR && ttv >= 0);
    const GridView& g = views[to.w];
    const VR_GLOBAL NanoLeaf* leaves = g.leaves;
    const float bg = g.background;
    const double vs = g.vs;
    int vrel;  // first vertex of this grid
    if constexpr (LIST) {
      vrel = ranges[4 * to.w];
    } else {
      const int g0 = tile_base[to.w];
      vrel = g0 < nt ? vbase[g0] : totals[1];
    }
    const int vbrel = vb - vrel;
    uint32_t rem = 0u;  // !LIST: cells not yet listed (mc_round_list)
    int rk = 0, n;
    if constexpr (LIST) {
      if (lane < VR_NBR) {
        el.tleaf[lane] = tl;
        el.ttile[lane] = ttv;
        el.self[lane] = (uint16_t)(cm & 0xFFFFu);
        el.fitem[lane] = (uint16_t)(cm >> 16);
      }
      n = mn;
    } else {
      const uint32_t mixed = mc_tile_rows<MASKS, true>(leaves, cq, bg, iso, tl, ttv, cm, el);
      const int mcnt = __popc(mixed);
      const int mincl = wave_incl_scan(mcnt);
      rem = mixed;
      rk = mincl - mcnt;
      n = __builtin_amdgcn_readlane(mincl, 63);
      if (MASKS && tn < t_end)
        q = mc_mask_issue(views[__builtin_amdgcn_readlane(h_tl, VR_NBR + 7)].signs, h_tl, lane);
    }
    const MCForeign fc{processed, s_ownsort, s_cand, el.ttile, meta, vbase, vrel};
    int vcarry = 0, tcarry = 0, fcarry = 0;
    // CPL mixed cells per lane and round, cell base + CPL * lane + j: LIST
    // takes rounds of 128 cells (two consecutive cells per lane), which
    // halves the rounds of a typical tile and fills the vertex and triangle
    // passes better; else rounds of 64.
    constexpr int CPL = LIST ? 2 : 1, RC = 64 * CPL;
    constexpr int RV = LIST ? MC_RVL : MC_RV, RT = LIST ? MC_RTL : MC_RT, NPC = LIST ? 3 : 1;
    if constexpr (LIST) wave_lds_sync();  // the tile's lists in LDS before the first round reads them
    for (int base = 0; base < n; base += RC) {
      uint32_t ent[CPL];
      if constexpr (LIST) {
        // this round's entries, and the loads of the next round's
        ent[0] = c0;
        ent[CPL - 1] = c1;
        c0 = c1 = 0u;
        const int nx = base + RC + 2 * lane;
        if (nx < n) c0 = mcell[(size_t)moff + nx];
        if (nx + 1 < n) c1 = mcell[(size_t)moff + nx + 1];
      } else {
        ent[0] = 0u;
        mc_round_list(&rem, &rk, base, lane, el.rlist);
        wave_lds_sync();
      }
      int cl[CPL], ntr[CPL];
      uint32_t owned[CPL], fm[CPL], cube[CPL];
      int pv = 0, pt = 0;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        cl[j] = 0;
        ntr[j] = 0;
        owned[j] = fm[j] = cube[j] = 0u;
        if (base + CPL * lane + j < n) {
          cl[j] = LIST ? (int)(ent[j] & 511u) : el.rlist[lane];
          const int ccls = s_ccls[cl[j]];
          cube[j] = LIST ? (ent[j] >> 9) : mc_case_rows(el.rows, cl[j] >> 6, (cl[j] >> 3) & 7, cl[j] & 7);
          if (!LIST) el.rcube[lane] = (uint8_t)cube[j];
          const uint32_t emn = s_emn[cube[j]];
          owned[j] = (emn & 0xFFFu) & el.self[ccls];
          fm[j] = (emn & 0xFFFu) & el.fitem[ccls];
          ntr[j] = (int)(emn >> 12);
        }
        pv += __popc(owned[j]);
        pt += ntr[j];
      }
      const int packed = pv | (pt << 16);
      const int incl = wave_incl_scan(packed);
      const int excl = incl - packed, tot = __builtin_amdgcn_readlane(incl, 63);
      const int rv = tot & 0xFFFF, rt = tot >> 16;
      const int vr = excl & 0xFFFF, tr = excl >> 16;  // the lane's first vertex and triangle in the round
      const bool listed = rv <= RV && rt <= RT;       // uniform
      if (listed) {
        int k = vr;
#pragma unroll
        for (int j = 0; j < CPL; ++j)
          for (uint32_t m = owned[j]; m; m &= m - 1u, ++k) el.vlist[k] = (uint16_t)((cl[j] << 4) | (__ffs(m) - 1));
      }
      // LIST: the corner values of the round's vertices (at most RV = 3 x
      // 64) are requested now and consumed after the foreign edges and the
      // triangle list are done.
      // (values in pva, pvb; the presence bits of vertex j at bits 2 j, 2 j + 1 of pok)
      float pva[NPC], pvb[NPC];
      uint32_t pok = 0u;
#pragma unroll
      for (int j = 0; j < NPC; ++j) pva[j] = pvb[j] = 0.0f;
      if (LIST && listed) {
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < NPC; ++j)
          if (lane + 64 * j < rv) {
            const int it = el.vlist[lane + 64 * j];
            const MCCorner c =
                mc_corner_load(leaves, el.tleaf, (it >> 10) & 7, (it >> 7) & 7, (it >> 4) & 7, it & 15, s_einfo);
            pva[j] = c.va;
            pvb[j] = c.vb;
            pok |= (uint32_t)c.ok << (2 * j);
          }
      }
      // Edges owned in an earlier tile, resolved after the corner values are
      // requested, so that the wait for the earlier tiles' face meta covers
      // both (a round of 128 cells holds a typical tile's cells: one wait
      // per tile).
      mc_foreign_pass(fm[0], cl[0], CPL > 1 ? fm[CPL - 1] : 0u, CPL > 1 ? cl[CPL - 1] : 0, lane,
                      LIST ? el.flist32 : el.flist, &fcarry, el, s_eslot, fc);
      if (listed) {
        // (the triangle list shares its LDS with the foreign list above)
        {
          int k = tr;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            if constexpr (LIST) {
              const uint64_t row = s_trin[cube[j]];
              for (int q2 = 0; q2 < ntr[j]; ++q2, ++k)
                el.tlist32[k] = (uint32_t)cl[j] | (((uint32_t)(row >> (12 * q2)) & 0xFFFu) << 9);
            } else {
              for (int q2 = 0; q2 < ntr[j]; ++q2, ++k) el.tlist[k] = (uint16_t)(lane | (q2 << 6));
            }
          }
        }
        wave_lds_sync();
        if constexpr (LIST) {
          // at most three vertices per lane, with their corner values requested above
#pragma unroll
          for (int j = 0; j < NPC; ++j) {
            const int k = lane + 64 * j;
            if (k < rv) {
              const int vi = vb + vcarry + k;
              const int it = el.vlist[k], cv = it >> 4, ve = it & 15;
              const int ux = cv >> 6, uy = (cv >> 3) & 7, uz = cv & 7;
              mc_edge_at(el.edge, mc_edge_off(ux, uy, uz, s_eslot[ve])) = (uint16_t)(vcarry + k);
              if (vi < vmax)
                mc_vertex_store(MCCorner{pva[j], pvb[j], (int)((pok >> (2 * j)) & 3u)}, bg, ux, uy, uz, ve, to, vs,
                                iso, s_einfo, verts + 3 * (size_t)vi);
            }
          }
        } else {
          for (int k = lane; k < rv; k += 64) {
            const int vi = vb + vcarry + k;
            const int it = el.vlist[k], cv = it >> 4, ve = it & 15;
            const int ux = cv >> 6, uy = (cv >> 3) & 7, uz = cv & 7;
            mc_edge_at(el.edge, mc_edge_off(ux, uy, uz, s_eslot[ve])) = (uint16_t)(vcarry + k);
            if (vi < vmax)
              mc_write_vertex(leaves, bg, el.tleaf, ux, uy, uz, ve, to, vs, iso, s_einfo, verts + 3 * (size_t)vi);
          }
        }
        wave_lds_sync();
        for (int k = lane; k < rt; k += 64) {
          const int ti = tb + tcarry + k;
          int ux, uy, uz;
          uint32_t tri_e;  // the triangle's three edges, 4 bits each
          if constexpr (LIST) {
            const uint32_t it = el.tlist32[k];
            ux = (int)((it >> 6) & 7u);
            uy = (int)((it >> 3) & 7u);
            uz = (int)(it & 7u);
            tri_e = it >> 9;
          } else {
            const uint32_t it = el.tlist[k];
            const int rc = (int)(it & 63u), q2 = (int)(it >> 6);
            const int cv = el.rlist[rc];
            ux = cv >> 6;
            uy = (cv >> 3) & 7;
            uz = cv & 7;
            tri_e = (uint32_t)(s_trin[el.rcube[rc]] >> (12 * q2));
          }
          // the three corners' table reads go out together; a corner on an
          // edge past the side table (0xFFFF) is resolved again after them
          const int e0 = (int)(tri_e & 15u), e1 = (int)((tri_e >> 4) & 15u), e2 = (int)((tri_e >> 8) & 15u);
          const uint32_t s0 = s_eslot[e0], s1 = s_eslot[e1], s2 = s_eslot[e2];
          const uint32_t v0 = mc_edge_get(el.edge, mc_edge_off(ux, uy, uz, s0)),
                         v1 = mc_edge_get(el.edge, mc_edge_off(ux, uy, uz, s1)),
                         v2 = mc_edge_get(el.edge, mc_edge_off(ux, uy, uz, s2));
          int i0 = mc_corner_known(v0, el, vbrel), i1 = mc_corner_known(v1, el, vbrel),
              i2 = mc_corner_known(v2, el, vbrel);
          if (v0 == 0xFFFFu || v1 == 0xFFFFu || v2 == 0xFFFFu) {
            if (v0 == 0xFFFFu) i0 = mc_foreign_index(ux, uy, uz, e0, fc.processed, fc.s_ownsort, fc.s_cand, fc.ttile,
                                                     fc.meta, fc.vbase, fc.vrel);
            if (v1 == 0xFFFFu) i1 = mc_foreign_index(ux, uy, uz, e1, fc.processed, fc.s_ownsort, fc.s_cand, fc.ttile,
                                                     fc.meta, fc.vbase, fc.vrel);
            if (v2 == 0xFFFFu) i2 = mc_foreign_index(ux, uy, uz, e2, fc.processed, fc.s_ownsort, fc.s_cand, fc.ttile,
                                                     fc.meta, fc.vbase, fc.vrel);
          }
          if (ti < tmax) {
            uint32_t* o = tris + 3 * (size_t)ti;
            o[0] = (uint32_t)i0;
            o[1] = (uint32_t)i1;
            o[2] = (uint32_t)i2;
          }
        }
      } else {
        // Dense round: cell by cell, every vertex (and its table slot)
        // before any triangle of the round.
        wave_lds_sync();
        {
          int vk = vcarry + vr;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const int cx = cl[j] >> 6, cy = (cl[j] >> 3) & 7, cz = cl[j] & 7;
            for (uint32_t m = owned[j]; m; m &= m - 1u, ++vk) {
              const int ve = __ffs(m) - 1;
              mc_edge_at(el.edge, mc_edge_off(cx, cy, cz, s_eslot[ve])) = (uint16_t)vk;
              if (vb + vk < vmax)
                mc_write_vertex(leaves, bg, el.tleaf, cx, cy, cz, ve, to, vs, iso, s_einfo,
                                verts + 3 * (size_t)(vb + vk));
            }
          }
        }
        wave_lds_sync();
        {
          int ti = tb + tcarry + tr;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const int cx = cl[j] >> 6, cy = (cl[j] >> 3) & 7, cz = cl[j] & 7;
            const uint64_t row = s_trin[cube[j]];
            for (int q2 = 0; q2 < ntr[j]; ++q2, ++ti)
              if (ti < tmax)
                for (int jj = 0; jj < 3; ++jj)
                  tris[3 * (size_t)ti + jj] = (uint32_t)mc_corner_index(
                      cx, cy, cz, (int)((row >> (4 * (3 * q2 + jj))) & 15u), el, s_eslot, vbrel, fc);
          }
        }
      }
      wave_lds_sync();
      vcarry += rv;
      tcarry += rt;
    }
  }
}


// ------------------------------------------------------------------ host driver
static int blocks_for(size_t n, int b) {
  size_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// Blocks that loop over the tiles (tables are loaded into LDS once per
// block, MC_WAVES tiles in flight per block): as many as are resident at
// once, since the tiles are split statically between them.
static int mc_blocks(const vr_ctx* ctx, const void* kernel, int waves, int tcap) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * waves, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const int cap = per_cu * (ctx->cus > 0 ? ctx->cus : 256);
  const int need = (tcap + waves - 1) / waves;
  return need < 1 ? 1 : (need < cap ? need : cap);
}

// Phase A: tile list for G grids (leaf counts known on the host: ltot).
// Leaves the tile count in s.totals[0]; caller reads nt.  single (G = 1):
// the grid's view on the host; the keep launch then writes d_views[0] and
// d_leaf_base[0..1] itself.
int mc_tiles_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot,
                   const GridView* single) {
  hipError_t e;
  const size_t n7 = (size_t)(7 * ltot);
  if ((e = s.keep.ensure(sizeof(int) * (n7 + 1))) != hipSuccess || (e = s.pos.ensure(sizeof(int) * (n7 + 1))) != hipSuccess ||
      (e = s.sums.ensure(sizeof(int) * (vr_scan_scratch_ints(n7) + 1))) != hipSuccess ||
      (e = s.totals.ensure(sizeof(int) * 4)) != hipSuccess ||
      (e = s.tile_base.ensure(sizeof(int) * (size_t)(G + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if (single)
    hipLaunchKernelGGL(k_mc_keep<true>, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views, d_leaf_base, G,
                       s.keep.as<int>(), (long long)n7, *single, const_cast<GridView*>(d_views),
                       const_cast<int*>(d_leaf_base));
  else
    hipLaunchKernelGGL(k_mc_keep<false>, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views, d_leaf_base, G,
                       s.keep.as<int>(), (long long)n7, GridView{}, (GridView*)nullptr, (int*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = vr_exclusive_scan(s.keep.as<int>(), s.pos.as<int>(), n7, s.sums.as<int>(), s.totals.as<int>(),
                             ctx->stream, d_leaf_base + G, 7)) != hipSuccess)
    return vr_hip_fail(ctx, e, "mc tiles");
  return VR_OK;
}

// Phase B (tile capacity tcap known): scatter (and tile_base), neighbours.
// Depends on the leaf sets only.
int mc_tiles_finish(vr_ctx* ctx, MCMulti& s, const GridView* d_views, const int* d_leaf_base, int G, long long ltot,
                    int tcap) {
  hipError_t e;
  const size_t n7 = (size_t)(7 * ltot);
  const size_t tc = (size_t)(tcap + 1);
  if ((e = s.tiles.ensure(sizeof(int4) * tc)) != hipSuccess || (e = s.tanchor.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.tleaf.ensure(sizeof(int) * VR_NBR * tc)) != hipSuccess ||
      (e = s.ttile.ensure(sizeof(int) * VR_NBR * tc)) != hipSuccess ||
      (e = s.cmask.ensure(sizeof(uint32_t) * VR_NBR * tc)) != hipSuccess ||
      (e = s.meta.ensure(sizeof(uint32_t) * MC_FACE * tc)) != hipSuccess || (e = s.tv.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.tt.ensure(sizeof(int) * tc)) != hipSuccess || (e = s.vb.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.tb.ensure(sizeof(int) * tc)) != hipSuccess ||
      (e = s.sums2.ensure(sizeof(int) * (vr_scan_scratch_ints(tc) + 1))) != hipSuccess ||
      (e = s.ranges.ensure(sizeof(int) * 4 * (size_t)(G + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  int* totals = s.totals.as<int>();
  if (G == 1 && ltot <= MC_SCATTER_LDS)
    hipLaunchKernelGGL(k_mc_scatter_tiles<true>, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views,
                       d_leaf_base, G, s.keep.as<int>(), s.pos.as<int>(), totals, s.tiles.as<int4>(),
                       s.tanchor.as<int>(), s.tile_base.as<int>());
  else
    hipLaunchKernelGGL(k_mc_scatter_tiles<false>, dim3(blocks_for(n7, 256)), dim3(256), 0, ctx->stream, d_views,
                       d_leaf_base, G, s.keep.as<int>(), s.pos.as<int>(), totals, s.tiles.as<int4>(),
                       s.tanchor.as<int>(), s.tile_base.as<int>());
  if (tcap >= MC_TN_BIG)
    hipLaunchKernelGGL(k_mc_tile_nbr, dim3(blocks_for((size_t)tcap, MC_TN_TILES)), dim3(256), 0, ctx->stream, d_views,
                       s.tiles.as<int4>(), s.tanchor.as<int>(), totals, s.tile_base.as<int>(), s.tleaf.as<int>(),
                       s.ttile.as<int>(), s.cmask.as<uint32_t>());
  else
    hipLaunchKernelGGL(k_mc_tile_nbr_slots, dim3(blocks_for((size_t)tcap, MC_TNS_TILES)), dim3(MC_TNS_TILES * VR_NBR), 0,
                       ctx->stream, d_views, s.tiles.as<int4>(), s.tanchor.as<int>(), totals, s.tile_base.as<int>(),
                       s.tleaf.as<int>(), s.ttile.as<int>(), s.cmask.as<uint32_t>());
  return (e = hipGetLastError()) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "mc tiles");
}

// Phase C: count, scans (the tile lists of phases A and B in s).
int mc_count_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, long long ltot, int tcap, float iso, bool masks,
                   long long* waves_out) {
  hipError_t e;
  int* totals = s.totals.as<int>();
  long long mcap = 0;
  if (masks) {
    if ((e = s.mtile.ensure(sizeof(int) * 2 * (size_t)(tcap + 1))) != hipSuccess ||
        (e = s.mctr.ensure(sizeof(unsigned long long))) != hipSuccess ||
        (e = vr_fill_int(s.mctr.as<int>(), 0, 2, ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "scratch");
    mcap = (long long)(s.mcell.bytes / sizeof(uint32_t));
    if (mcap > 0x7fffffffLL) mcap = 0x7fffffffLL;  // offsets are stored as int
  }
  {
    TimedLaunch tl(ctx, TC_MC, masks ? 64.0 : (double)VR_LEAF_BYTES, 0.0, (double)ltot);
    const void* k = masks ? (const void*)k_mc_count<true, true> : (const void*)k_mc_count<false, false>;
    const int blocks = mc_blocks(ctx, k, MC_WAVES, tcap);
    if (waves_out) *waves_out = (long long)blocks * MC_WAVES;
    hipLaunchKernelGGL((masks ? k_mc_count<true, true> : k_mc_count<false, false>),
                       dim3(blocks), dim3(MC_THREADS), 0, ctx->stream, d_views,
                       s.tiles.as<int4>(), totals, s.tleaf.as<int>(), s.cmask.as<uint32_t>(), iso,
                       s.meta.as<uint32_t>(), s.tv.as<int>(), s.tt.as<int>(), tcap, s.mcell.as<uint32_t>(),
                       s.mtile.as<int>(), s.mctr.as<unsigned long long>(), mcap);
  }
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = vr_exclusive_scan2(s.tv.as<int>(), s.vb.as<int>(), totals + 1, s.tt.as<int>(), s.tb.as<int>(),
                              totals + 2, (size_t)tcap, s.sums2.as<int>(), ctx->stream, totals)) != hipSuccess)
    return vr_hip_fail(ctx, e, "mc count");
  return VR_OK;
}

// Phase D: emit into verts/tris (capacities from the counts), and the
// per-grid ranges (s.ranges; counts_out, optional: the totals).
int mc_emit_multi(vr_ctx* ctx, MCMulti& s, const GridView* d_views, int G, int tcap, float iso, double* verts,
                  uint32_t* tris, double leaf_count, double mesh_bytes, long long vcap, long long tricap,
                  int* counts_out, bool masks) {
  hipError_t e;
  if (vcap < 0) vcap = 0x7fffffffLL;
  if (tricap < 0) tricap = 0x7fffffffLL;
  {
    // with masks (cells from the count pass's lists): each leaf's 2048 value bytes (vertices interpolate values
    // throughout the leaf); the count pass reads the 64-byte sign masks
    TimedLaunch tl(ctx, TC_MC_EMIT, masks ? 2048.0 : (double)VR_LEAF_BYTES, mesh_bytes, leaf_count, true);
    if (masks)
      hipLaunchKernelGGL(k_mc_ranges, dim3(1), dim3(256), 0, ctx->stream, s.tile_base.as<int>(), G, s.totals.as<int>(),
                         s.vb.as<int>(), s.tb.as<int>(), s.totals.as<int>(), s.ranges.as<int>(), counts_out);
    const void* k = masks ? (const void*)k_mc_emit<true, true> : (const void*)k_mc_emit<false, false>;
    hipLaunchKernelGGL((masks ? k_mc_emit<true, true> : k_mc_emit<false, false>),
                       dim3(mc_blocks(ctx, k, MC_EWAVES, tcap)), dim3(MC_ETHREADS), 0, ctx->stream, d_views,
                       s.tiles.as<int4>(), s.totals.as<int>(), s.tleaf.as<int>(), s.ttile.as<int>(),
                       s.tile_base.as<int>(), s.cmask.as<uint32_t>(), iso, s.meta.as<uint32_t>(), s.tv.as<int>(),
                       s.tt.as<int>(), s.vb.as<int>(), s.tb.as<int>(), s.totals.as<int>(), verts, tris, vcap, tricap,
                       G, s.ranges.as<int>(), counts_out, s.mcell.as<uint32_t>(), s.mtile.as<int>());
  }
  return (e = hipGetLastError()) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "mc emit");
}

// Single grid (volume_to_mesh) through the multi-grid passes with G = 1.
int mc_run(vr_ctx* ctx, const vr_grid* g, float iso, MCScratch* unused, vr_mesh* m, bool deferred, int* d_counts) {
  (void)unused;
  hipError_t e;
  const int n = g->topo->count;  // exact, or an upper bound when deferred
  MCMulti& s = ctx->mc;
  if ((e = s.views.ensure(sizeof(GridView))) != hipSuccess || (e = s.leaf_base.ensure(sizeof(int) * 2)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  // the view and the leaf range [0, count) go to the device with the first pass
  const GridView v = make_view(g);
  int rc = mc_tiles_multi(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, &v);
  if (rc != VR_OK) return rc;
  if (deferred) {
    // Capacities from the bounds: at most 7 tiles per leaf; the mesh keeps
    // the capacity it has (a first call takes a guess); the counts go to
    // d_counts and the caller checks them against the capacity.
    const int tcap = 7 * (n < 1 ? 1 : n);
    rc = mc_tiles_finish(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, tcap);
    if (rc == VR_OK) rc = mc_count_multi(ctx, s, s.views.as<GridView>(), n, tcap, iso);
    if (rc != VR_OK) return rc;
    if (m->cap_v < 1 || m->cap_t < 1) {
      const uint64_t cv = (uint64_t)16 * (uint64_t)(n < 256 ? 256 : n);
      if (m->d_v) (void)hipFree(m->d_v);
      if (m->d_t) (void)hipFree(m->d_t);
      m->d_v = nullptr;
      m->d_t = nullptr;
      m->cap_v = m->cap_t = 0;
      if ((e = hipMalloc(&m->d_v, sizeof(double) * 3 * cv)) != hipSuccess ||
          (e = hipMalloc(&m->d_t, sizeof(uint32_t) * 3 * 2 * cv)) != hipSuccess)
        return vr_hip_fail(ctx, e, "hipMalloc mesh");
      m->cap_v = cv;
      m->cap_t = 2 * cv;
    }
    return mc_emit_multi(ctx, s, s.views.as<GridView>(), 1, tcap, iso, m->d_v, m->d_t, (double)n, 0.0,
                         (long long)m->cap_v,
                         (long long)m->cap_t, d_counts);
  }
  int nt = 0;
  if ((e = hipMemcpyAsync(&nt, s.totals.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "readback");
  rc = mc_tiles_finish(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, nt);
  if (rc == VR_OK) rc = mc_count_multi(ctx, s, s.views.as<GridView>(), n, nt, iso);
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
  return mc_emit_multi(ctx, s, s.views.as<GridView>(), 1, nt, iso, m->d_v, m->d_t, (double)n,
                       24.0 * th[1] + 12.0 * th[2]);
}

// ------------------------------------------------------------------ allocation-free entry point
// vr_marching_cubes_noalloc keeps the passes' scratch (tile lists,
// neighbour slots, face meta, per-tile counts, scan partials) in one
// file-scope device array laid out for grids of up to VR_NOALLOC_MAX_LEAVES
// leaves (7 tiles per leaf at most), instead of context buffers that grow
// with hipMalloc.  Each device holds its own copy until the process ends;
// its address comes from hipGetSymbolAddress, taken once per device.
#define VR_NOALLOC_ALIGN 256

// Byte offsets of the MCMulti ranges for n leaves (tile capacity 7 n); the
// same function sizes the array and carves it.
struct MCNoAllocLayout {
  size_t off[19];
  size_t len[19];
  size_t total;
};
static constexpr size_t vr_align_up(size_t x) { return (x + VR_NOALLOC_ALIGN - 1) / VR_NOALLOC_ALIGN * VR_NOALLOC_ALIGN; }
static constexpr size_t vr_scan_ints(size_t n) { return (n + 1023) / 1024 + 1; }  // vr_scan_scratch_ints
static constexpr MCNoAllocLayout mc_noalloc_layout(size_t n) {
  MCNoAllocLayout L{};
  const size_t n7 = 7 * n, tc = 7 * (n < 1 ? 1 : n) + 1;
  const size_t len[19] = {sizeof(GridView), 2 * sizeof(int), sizeof(int) * (n7 + 1), sizeof(int) * (n7 + 1),
                          sizeof(int) * (vr_scan_ints(n7) + 1), sizeof(int) * (vr_scan_ints(tc) + 1), 4 * sizeof(int),
                          2 * sizeof(int), 16 * tc, sizeof(int) * tc, sizeof(int) * VR_NBR * tc,
                          sizeof(int) * VR_NBR * tc, sizeof(uint32_t) * VR_NBR * tc, sizeof(uint32_t) * MC_FACE * tc,
                          sizeof(int) * tc, sizeof(int) * tc, sizeof(int) * tc, sizeof(int) * tc, 8 * sizeof(int)};
  size_t at = 0;
  for (int i = 0; i < 19; ++i) {
    L.off[i] = at;
    L.len[i] = vr_align_up(len[i]);
    at += L.len[i];
  }
  L.total = at;
  return L;
}
static constexpr size_t VR_NOALLOC_BYTES = mc_noalloc_layout(VR_NOALLOC_MAX_LEAVES).total;

static __device__ __align__(256) unsigned char scratch[VR_NOALLOC_BYTES];

#define VR_NOALLOC_DEVICES 64
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[VR_NOALLOC_DEVICES] = {};

hipError_t mc_noalloc_setup(int device) {
  if (device < 0 || device >= VR_NOALLOC_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_noalloc_ptr[device] = ptr;
  return hipSuccess;
}

static void* mc_noalloc_ptr(int device) {
  if (device < 0 || device >= VR_NOALLOC_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

extern "C" int vr_marching_cubes_noalloc(vr_ctx* ctx, const vr_grid* g, float iso, double* xyz, uint64_t vcap,
                                         uint32_t* tri, uint64_t tcap, int32_t* counts) {
  if (!ctx || !g || !counts || (vcap && !xyz) || (tcap && !tri))
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  const int n = g->topo->count;
  if (n > VR_NOALLOC_MAX_LEAVES)
    return vr_fail(ctx, VR_ERR_CAPACITY, "grid has %d leaves; the allocation-free marching cubes takes at most %d", n,
                   VR_NOALLOC_MAX_LEAVES);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  unsigned char* base = static_cast<unsigned char*>(mc_noalloc_ptr(ctx->device));
  if (!base) return vr_fail(ctx, VR_ERR_HIP, "no scratch address for device %d", ctx->device);
  if (n == 0) {  // no tiles: the counts are zero (a fill kernel, no memset node)
    e = vr_fill_int(counts, 0, 2, ctx->stream);
    return e == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "fill");
  }
  // The passes' buffers as borrowed ranges of this device's array, laid out
  // for this grid's leaf count (every range within the array: n <= max).
  constexpr int kSlots = 19;
  const MCNoAllocLayout L = mc_noalloc_layout((size_t)n);
  MCMulti s;
  DevBuf* slot[kSlots] = {&s.views, &s.leaf_base, &s.keep, &s.pos, &s.sums, &s.sums2, &s.totals, &s.tile_base,
                          &s.tiles, &s.tanchor, &s.tleaf, &s.ttile, &s.cmask, &s.meta, &s.tv, &s.tt, &s.vb, &s.tb,
                          &s.ranges};
  for (int i = 0; i < kSlots; ++i) slot[i]->borrow(base + L.off[i], L.len[i]);
  // No event records (they would create events) inside this call.
  const bool timing = ctx->timing;
  ctx->timing = false;
  const GridView v = make_view(g);
  const int tc = 7 * n;
  int rc = mc_tiles_multi(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, &v);
  if (rc == VR_OK) rc = mc_tiles_finish(ctx, s, s.views.as<GridView>(), s.leaf_base.as<int>(), 1, n, tc);
  if (rc == VR_OK) rc = mc_count_multi(ctx, s, s.views.as<GridView>(), n, tc, iso);
  if (rc == VR_OK)
    rc = mc_emit_multi(ctx, s, s.views.as<GridView>(), 1, tc, iso, xyz, tri, (double)n, 0.0,
                       (long long)(vcap < 0x7fffffffull ? vcap : 0x7fffffffull),
                       (long long)(tcap < 0x7fffffffull ? tcap : 0x7fffffffull), counts);
  ctx->timing = timing;
  return rc;
}
