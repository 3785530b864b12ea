// adaptive_surface (src/tools/adaptive_mesh.rs:100-241) on the device.
//
// Marching cubes gives the base mesh (vr_mc.hip).  Each subdivision depth:
//   1. every vertex's |mean curvature| at its rounded index position
//      (gradient.rs:83-135 through tree lookups, adaptive_mesh.rs:47-55);
//   2. every triangle's split test (max vertex curvature above the
//      threshold and longest edge above max_edge_length / 2^(depth+1)); a
//      split triangle enters its three edges, keyed (min, max) vertex, into
//      a hash table with the least position 3 t + s that names the edge
//      (s = 0, 1, 2 for edges 01, 12, 20);
//   3. a position holding its edge's least position is the edge's first
//      encounter; an exclusive scan over positions numbers the new vertices
//      in the reference's order (triangle order, then 01, 12, 20), so
//      vertex V + rank is the midpoint of that edge projected onto the
//      isosurface (adaptive_mesh.rs:70-78, 243-259);
//   4. each triangle writes itself or its four children at the scanned
//      offset, in the reference's child order.
// Normals: every vertex gathers its (triangle, corner) incidences, puts them
// in triangle order and sums the angle-weighted face normals in that order,
// as the reference's single loop does (adaptive_mesh.rs:166-222).  Signed
// curvatures at the end (:225-229).  Floating-point operations follow the
// reference's order and types (f32 differences, f64 geometry, no
// contraction); acos is the device's.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "vr_internal.h"

struct AdaptParams {
  double inv[12];  // inverse affine rows 0..2 (grid.rs:268-277)
  double vs;       // affine voxel size (grid.rs:281-283)
  float iso;
};

// AffineMap::world_to_index (grid.rs:261-264): f64::round (half away from
// zero), then `as i32`.
__device__ __forceinline__ void ad_world_to_index(const AdaptParams& p, const double w[3], int c[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double f = p.inv[4 * r] * w[0] + p.inv[4 * r + 1] * w[1] + p.inv[4 * r + 2] * w[2] + p.inv[4 * r + 3];
    c[r] = f64_to_i32(round(f));
  }
}

// gradient.rs:83-135, in its association order.
__device__ float ad_mean_curvature(const GridView& g, int x, int y, int z) {
  const float xp = grid_get(g, x + 1, y, z), xm = grid_get(g, x - 1, y, z);
  const float yp = grid_get(g, x, y + 1, z), ym = grid_get(g, x, y - 1, z);
  const float zp = grid_get(g, x, y, z + 1), zm = grid_get(g, x, y, z - 1);
  const float phi_x = (xp - xm) * 0.5f, phi_y = (yp - ym) * 0.5f, phi_z = (zp - zm) * 0.5f;
  const float center = grid_get(g, x, y, z);
  const float phi_xx = xp - 2.0f * center + xm;
  const float phi_yy = yp - 2.0f * center + ym;
  const float phi_zz = zp - 2.0f * center + zm;
  const float phi_xy = (grid_get(g, x + 1, y + 1, z) - grid_get(g, x + 1, y - 1, z) - grid_get(g, x - 1, y + 1, z) +
                        grid_get(g, x - 1, y - 1, z)) *
                       0.25f;
  const float phi_xz = (grid_get(g, x + 1, y, z + 1) - grid_get(g, x + 1, y, z - 1) - grid_get(g, x - 1, y, z + 1) +
                        grid_get(g, x - 1, y, z - 1)) *
                       0.25f;
  const float phi_yz = (grid_get(g, x, y + 1, z + 1) - grid_get(g, x, y + 1, z - 1) - grid_get(g, x, y - 1, z + 1) +
                        grid_get(g, x, y - 1, z - 1)) *
                       0.25f;
  const float grad_mag_sq = phi_x * phi_x + phi_y * phi_y + phi_z * phi_z;
  const float grad_mag = sqrtf(grad_mag_sq);
  if (grad_mag < 1e-10f) return 0.0f;
  const float numerator = phi_xx * (phi_y * phi_y + phi_z * phi_z) + phi_yy * (phi_x * phi_x + phi_z * phi_z) +
                          phi_zz * (phi_x * phi_x + phi_y * phi_y) -
                          2.0f * (phi_x * phi_y * phi_xy + phi_x * phi_z * phi_xz + phi_y * phi_z * phi_yz);
  return numerator / (grad_mag_sq * grad_mag);
}

// Per vertex: curvature_at_world (adaptive_mesh.rs:47-55), |.| when absval.
__global__ __launch_bounds__(256) void k_ad_curvature(GridView g, AdaptParams p, const double* v, int nv, double* out,
                                                      int absval) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    const double w[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    int c[3];
    ad_world_to_index(p, w, c);
    const double h = (double)ad_mean_curvature(g, c[0], c[1], c[2]) / p.vs;
    out[i] = absval ? fabs(h) : h;
  }
}

// f64::max: the other operand when one is NaN.
__device__ __forceinline__ double ad_max(double a, double b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }

__device__ __forceinline__ double ad_edge(const double* v, uint32_t a, uint32_t b) {
  const double dx = v[3 * b] - v[3 * a], dy = v[3 * b + 1] - v[3 * a + 1], dz = v[3 * b + 2] - v[3 * a + 2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

#define AD_EMPTY 0xFFFFFFFFFFFFFFFFull

__device__ __forceinline__ uint32_t ad_hash(uint64_t k) {
  k ^= k >> 33;
  k *= 0xFF51AFD7ED558CCDull;
  k ^= k >> 33;
  return (uint32_t)k;
}

// Split test per triangle (adaptive_mesh.rs:131-160); a split triangle's
// edges go into the table with their least position.
__global__ __launch_bounds__(256) void k_ad_split(const double* v, const uint32_t* tri, int nt, const double* vc,
                                                  double threshold, double min_edge, int* flag, uint64_t* hkey,
                                                  int* hpos, uint32_t hmask) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    const double max_curv = ad_max(ad_max(vc[a], vc[b]), vc[c]);
    const double max_edge = ad_max(ad_max(ad_edge(v, a, b), ad_edge(v, b, c)), ad_edge(v, c, a));
    const int s = max_curv > threshold && max_edge > min_edge ? 1 : 0;
    flag[t] = s;
    if (!s) continue;
    const uint32_t ends[4] = {a, b, c, a};
    for (int q = 0; q < 3; ++q) {
      const uint32_t lo = min(ends[q], ends[q + 1]), hi = max(ends[q], ends[q + 1]);
      const uint64_t key = ((uint64_t)lo << 32) | hi;
      uint32_t h = ad_hash(key) & hmask;
      while (true) {
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(hkey + h), AD_EMPTY, key);
        if (prev == AD_EMPTY || prev == key) {
          atomicMin(hpos + h, 3 * t + q);
          break;
        }
        h = (h + 1) & hmask;
      }
    }
  }
}

// Per position 3 t + s of a split triangle: its table slot and whether it is
// its edge's first encounter.
__global__ __launch_bounds__(256) void k_ad_first(const uint32_t* tri, const int* flag, int nt, const uint64_t* hkey,
                                                  const int* hpos, uint32_t hmask, int* first, int* eslot) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 3 * nt; q += gridDim.x * blockDim.x) {
    const int t = q / 3, s = q % 3;
    if (!flag[t]) {
      first[q] = 0;
      continue;
    }
    const uint32_t e0 = tri[3 * t + s], e1 = tri[3 * t + (s == 2 ? 0 : s + 1)];
    const uint64_t key = ((uint64_t)min(e0, e1) << 32) | max(e0, e1);
    uint32_t h = ad_hash(key) & hmask;
    while (hkey[h] != key) h = (h + 1) & hmask;  // present: inserted by k_ad_split
    eslot[q] = (int)h;
    first[q] = hpos[h] == q ? 1 : 0;
  }
}

// normal_at_world (adaptive_mesh.rs:58-67): the squared length summed in f32.
__device__ __forceinline__ void ad_normal(const GridView& g, const AdaptParams& p, const double w[3], double n[3]) {
  int c[3];
  ad_world_to_index(p, w, c);
  const float gx = (grid_get(g, c[0] + 1, c[1], c[2]) - grid_get(g, c[0] - 1, c[1], c[2])) * 0.5f;
  const float gy = (grid_get(g, c[0], c[1] + 1, c[2]) - grid_get(g, c[0], c[1] - 1, c[2])) * 0.5f;
  const float gz = (grid_get(g, c[0], c[1], c[2] + 1) - grid_get(g, c[0], c[1], c[2] - 1)) * 0.5f;
  const double len = sqrt((double)(gx * gx + gy * gy + gz * gz));
  if (len < 1e-15) {
    n[0] = 0.0;
    n[1] = 0.0;
    n[2] = 1.0;
  } else {
    n[0] = (double)gx / len;
    n[1] = (double)gy / len;
    n[2] = (double)gz / len;
  }
}

// New vertex V + rank at each first encounter: the edge's midpoint projected
// onto the isosurface; midx[slot] = its index.
__global__ __launch_bounds__(256) void k_ad_midpoints(GridView g, AdaptParams p, const uint32_t* tri, const int* first,
                                                      const int* rank, const int* eslot, int nt, int V, double* v,
                                                      int* midx) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 3 * nt; q += gridDim.x * blockDim.x) {
    if (!first[q]) continue;
    const int t = q / 3, s = q % 3;
    const uint32_t a = tri[3 * t + s], b = tri[3 * t + (s == 2 ? 0 : s + 1)];
    const double m[3] = {(v[3 * a] + v[3 * b]) * 0.5, (v[3 * a + 1] + v[3 * b + 1]) * 0.5,
                         (v[3 * a + 2] + v[3 * b + 2]) * 0.5};
    const float sdf = sample_grid(g, m[0], m[1], m[2]) - p.iso;
    double n[3];
    ad_normal(g, p, m, n);
    const int idx = V + rank[q];
    v[3 * (size_t)idx] = m[0] - (double)sdf * n[0];
    v[3 * (size_t)idx + 1] = m[1] - (double)sdf * n[1];
    v[3 * (size_t)idx + 2] = m[2] - (double)sdf * n[2];
    midx[eslot[q]] = idx;
  }
}

// The next depth's triangles (adaptive_mesh.rs:152-159): triangle t starts
// at t + 3 * (split triangles before t), the exclusive scan of the flags.
__global__ __launch_bounds__(256) void k_ad_tris(const uint32_t* tri, const int* flag, const int* eslot, const int* midx,
                                                 const int* nsplit_before, int nt, uint32_t* out) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    uint32_t* o = out + 3 * ((size_t)t + 3 * (size_t)nsplit_before[t]);
    if (!flag[t]) {
      o[0] = a;
      o[1] = b;
      o[2] = c;
      continue;
    }
    const uint32_t m01 = (uint32_t)midx[eslot[3 * t]], m12 = (uint32_t)midx[eslot[3 * t + 1]],
                   m20 = (uint32_t)midx[eslot[3 * t + 2]];
    const uint32_t w[12] = {a, m01, m20, b, m12, m01, c, m20, m12, m01, m12, m20};
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = w[k];
  }
}

__global__ __launch_bounds__(256) void k_ad_degree(const uint32_t* tri, int nt, int* deg) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 3 * nt; q += gridDim.x * blockDim.x) atomicAdd(deg + tri[q], 1);
}

__global__ __launch_bounds__(256) void k_ad_incidence(const uint32_t* tri, int nt, const int* off, int* cursor, int* ent) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 3 * nt; q += gridDim.x * blockDim.x) {
    const uint32_t v = tri[q];
    ent[off[v] + atomicAdd(cursor + v, 1)] = q;
  }
}

// adaptive_mesh.rs:261-271
__device__ __forceinline__ double ad_angle(const double a[3], const double b[3]) {
  const double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (la < 1e-15 || lb < 1e-15) return 0.0;
  double c = dot / (la * lb);
  c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
  return acos(c);
}

// Per vertex: its incidences (3 t + k) in ascending order, each adding the
// face normal of triangle t times the angle at corner k; then normalised.
__global__ __launch_bounds__(256) void k_ad_normals(const double* v, const uint32_t* tri, const int* off, const int* deg,
                                                    int* ent, int nv, double* normals) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    int* e = ent + off[i];
    const int d = deg[i];
    for (int a = 1; a < d; ++a)  // insertion sort: a vertex has few incidences
      for (int b = a; b > 0 && e[b] < e[b - 1]; --b) {
        const int tmp = e[b];
        e[b] = e[b - 1];
        e[b - 1] = tmp;
      }
    double nx = 0.0, ny = 0.0, nz = 0.0;
    for (int k = 0; k < d; ++k) {
      const int t = e[k] / 3, corner = e[k] % 3;
      const uint32_t i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      const double v0[3] = {v[3 * i0], v[3 * i0 + 1], v[3 * i0 + 2]};
      const double v1[3] = {v[3 * i1], v[3 * i1 + 1], v[3 * i1 + 2]};
      const double v2[3] = {v[3 * i2], v[3 * i2 + 1], v[3 * i2 + 2]};
      const double e01[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      const double e02[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      const double e12[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
      const double fx = e01[1] * e02[2] - e01[2] * e02[1];
      const double fy = e01[2] * e02[0] - e01[0] * e02[2];
      const double fz = e01[0] * e02[1] - e01[1] * e02[0];
      double ang;
      if (corner == 0) {
        ang = ad_angle(e01, e02);
      } else if (corner == 1) {
        const double m01[3] = {-e01[0], -e01[1], -e01[2]};
        ang = ad_angle(m01, e12);
      } else {
        const double m02[3] = {-e02[0], -e02[1], -e02[2]}, m12[3] = {-e12[0], -e12[1], -e12[2]};
        ang = ad_angle(m02, m12);
      }
      nx += fx * ang;
      ny += fy * ang;
      nz += fz * ang;
    }
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    if (len > 1e-15) {
      nx /= len;
      ny /= len;
      nz /= len;
    }
    normals[3 * (size_t)i] = nx;
    normals[3 * (size_t)i + 1] = ny;
    normals[3 * (size_t)i + 2] = nz;
  }
}

struct vr_render_mesh {
  uint64_t nv = 0, nt = 0;
  double* d_v = nullptr;
  double* d_n = nullptr;
  double* d_c = nullptr;
  uint32_t* d_t = nullptr;
  vr_render_mesh() = default;
  vr_render_mesh(const vr_render_mesh&) = delete;
  vr_render_mesh& operator=(const vr_render_mesh&) = delete;
  ~vr_render_mesh() {
    if (d_v) (void)hipFree(d_v);
    if (d_n) (void)hipFree(d_n);
    if (d_c) (void)hipFree(d_c);
    if (d_t) (void)hipFree(d_t);
  }
};

static int ad_blocks(size_t n) {
  const size_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : (int)b);
}

extern "C" int vr_adaptive_surface(vr_ctx* ctx, const vr_grid* g, float iso, double max_edge_length,
                                   double curvature_threshold, uint32_t max_depth, vr_render_mesh** out) {
  if (!ctx || !g || !out) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (max_depth > 30) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "max_depth above 30");
  *out = nullptr;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  std::unique_ptr<vr_render_mesh> rm(new vr_render_mesh());
  if (g->topo->count == 0) {  // volume_to_mesh of an empty grid is empty
    *out = rm.release();
    return VR_OK;
  }
  vr_mesh base;
  MCScratch sc;
  int rc = mc_run(ctx, g, iso, &sc, &base);
  if (rc != VR_OK) return rc;
  AdaptParams p;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) p.inv[4 * r + c] = g->inv[r][c];
  p.vs = g->voxel_size;
  p.iso = iso;
  const GridView gv = make_view(g);
  long long V = (long long)base.nv, T = (long long)base.nt;
  // vertices and triangles of the current depth (the base mesh's buffers first)
  DevBuf verts, tris, tris2, vc, flag, first, rank, eslot, midx, hkey, hpos, off, sums, tot;
  if ((e = verts.ensure(sizeof(double) * 3 * (size_t)(V + 1))) != hipSuccess ||
      (e = tris.ensure(sizeof(uint32_t) * 3 * (size_t)(T + 1))) != hipSuccess ||
      (e = tot.ensure(sizeof(int) * 4)) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if ((V && (e = hipMemcpyAsync(verts.p, base.d_v, sizeof(double) * 3 * V, hipMemcpyDeviceToDevice, ctx->stream)) !=
                hipSuccess) ||
      (T && (e = hipMemcpyAsync(tris.p, base.d_t, sizeof(uint32_t) * 3 * T, hipMemcpyDeviceToDevice, ctx->stream)) !=
                hipSuccess))
    return vr_hip_fail(ctx, e, "copy");
  for (uint32_t depth = 0; depth < max_depth && T > 0; ++depth) {
    if (3 * T > 0x3fffffffLL || V + 3 * T > 0x7fffffffLL) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "mesh too large");
    const double min_edge = max_edge_length / (double)(1ull << (depth + 1));
    uint32_t hsize = 1024;
    while (hsize < (uint32_t)(6 * T)) hsize <<= 1;
    if ((e = vc.ensure(sizeof(double) * V)) != hipSuccess || (e = flag.ensure(sizeof(int) * T)) != hipSuccess ||
        (e = first.ensure(sizeof(int) * 3 * T)) != hipSuccess || (e = rank.ensure(sizeof(int) * 3 * T)) != hipSuccess ||
        (e = eslot.ensure(sizeof(int) * 3 * T)) != hipSuccess || (e = midx.ensure(sizeof(int) * hsize)) != hipSuccess ||
        (e = hkey.ensure(sizeof(uint64_t) * hsize)) != hipSuccess || (e = hpos.ensure(sizeof(int) * hsize)) != hipSuccess ||
        (e = off.ensure(sizeof(int) * T)) != hipSuccess ||
        (e = sums.ensure(sizeof(int) * (vr_scan_scratch_ints((size_t)(3 * T)) + 1))) != hipSuccess)
      return vr_hip_fail(ctx, e, "scratch");
    if ((e = vr_fill_int(hkey.as<int>(), -1, 2 * (size_t)hsize, ctx->stream)) != hipSuccess ||
        (e = vr_fill_int(hpos.as<int>(), 0x7fffffff, hsize, ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "fill");
    hipLaunchKernelGGL(k_ad_curvature, dim3(ad_blocks(V)), dim3(256), 0, ctx->stream, gv, p, verts.as<double>(), (int)V,
                       vc.as<double>(), 1);
    hipLaunchKernelGGL(k_ad_split, dim3(ad_blocks(T)), dim3(256), 0, ctx->stream, verts.as<double>(),
                       tris.as<uint32_t>(), (int)T, vc.as<double>(), curvature_threshold, min_edge, flag.as<int>(),
                       hkey.as<uint64_t>(), hpos.as<int>(), hsize - 1);
    hipLaunchKernelGGL(k_ad_first, dim3(ad_blocks(3 * T)), dim3(256), 0, ctx->stream, tris.as<uint32_t>(),
                       flag.as<int>(), (int)T, hkey.as<uint64_t>(), hpos.as<int>(), hsize - 1, first.as<int>(),
                       eslot.as<int>());
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = vr_exclusive_scan(first.as<int>(), rank.as<int>(), (size_t)(3 * T), sums.as<int>(), tot.as<int>(),
                               ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "subdivide");
    int nnew = 0;
    if ((e = hipMemcpyAsync(&nnew, tot.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "readback");
    if (nnew > 0) {
      // the vertex array grows: copy the current vertices into a larger one
      DevBuf grown;
      if ((e = grown.ensure(sizeof(double) * 3 * (size_t)(V + nnew))) != hipSuccess ||
          (e = hipMemcpyAsync(grown.p, verts.p, sizeof(double) * 3 * V, hipMemcpyDeviceToDevice, ctx->stream)) !=
              hipSuccess)
        return vr_hip_fail(ctx, e, "grow");
      std::swap(grown.p, verts.p);
      std::swap(grown.bytes, verts.bytes);
    }
    hipLaunchKernelGGL(k_ad_midpoints, dim3(ad_blocks(3 * T)), dim3(256), 0, ctx->stream, gv, p, tris.as<uint32_t>(),
                       first.as<int>(), rank.as<int>(), eslot.as<int>(), (int)T, (int)V, verts.as<double>(),
                       midx.as<int>());
    // split triangles before each triangle (the scanned flags)
    if ((e = vr_exclusive_scan(flag.as<int>(), off.as<int>(), (size_t)T, sums.as<int>(), tot.as<int>() + 1,
                               ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "scan");
    int nsplit = 0;
    if ((e = hipMemcpyAsync(&nsplit, tot.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "readback");
    const long long T2 = T + 3LL * nsplit;
    if ((e = tris2.ensure(sizeof(uint32_t) * 3 * (size_t)(T2 + 1))) != hipSuccess) return vr_hip_fail(ctx, e, "scratch");
    hipLaunchKernelGGL(k_ad_tris, dim3(ad_blocks(T)), dim3(256), 0, ctx->stream, tris.as<uint32_t>(), flag.as<int>(),
                       eslot.as<int>(), midx.as<int>(), off.as<int>(), (int)T, tris2.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return vr_hip_fail(ctx, e, "subdivide");
    std::swap(tris.p, tris2.p);
    std::swap(tris.bytes, tris2.bytes);
    V += nnew;
    T = T2;
  }
  // normals and signed curvatures of the final mesh
  if ((e = hipMalloc(&rm->d_v, sizeof(double) * 3 * (size_t)(V + 1))) != hipSuccess ||
      (e = hipMalloc(&rm->d_n, sizeof(double) * 3 * (size_t)(V + 1))) != hipSuccess ||
      (e = hipMalloc(&rm->d_c, sizeof(double) * (size_t)(V + 1))) != hipSuccess ||
      (e = hipMalloc(&rm->d_t, sizeof(uint32_t) * 3 * (size_t)(T + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "hipMalloc");
  DevBuf deg, voff, cursor, ent, vsums;
  if ((e = deg.ensure(sizeof(int) * (V + 1))) != hipSuccess || (e = voff.ensure(sizeof(int) * (V + 1))) != hipSuccess ||
      (e = cursor.ensure(sizeof(int) * (V + 1))) != hipSuccess ||
      (e = ent.ensure(sizeof(int) * 3 * (size_t)(T + 1))) != hipSuccess ||
      (e = vsums.ensure(sizeof(int) * (vr_scan_scratch_ints((size_t)V) + 1))) != hipSuccess)
    return vr_hip_fail(ctx, e, "scratch");
  if ((e = vr_fill_int(deg.as<int>(), 0, (size_t)V, ctx->stream)) != hipSuccess ||
      (e = vr_fill_int(cursor.as<int>(), 0, (size_t)V, ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "fill");
  if (T > 0) {
    hipLaunchKernelGGL(k_ad_degree, dim3(ad_blocks(3 * T)), dim3(256), 0, ctx->stream, tris.as<uint32_t>(), (int)T,
                       deg.as<int>());
    if ((e = vr_exclusive_scan(deg.as<int>(), voff.as<int>(), (size_t)V, vsums.as<int>(), tot.as<int>() + 2,
                               ctx->stream)) != hipSuccess)
      return vr_hip_fail(ctx, e, "scan");
    hipLaunchKernelGGL(k_ad_incidence, dim3(ad_blocks(3 * T)), dim3(256), 0, ctx->stream, tris.as<uint32_t>(), (int)T,
                       voff.as<int>(), cursor.as<int>(), ent.as<int>());
  } else if ((e = vr_fill_int(voff.as<int>(), 0, (size_t)V, ctx->stream)) != hipSuccess) {
    return vr_hip_fail(ctx, e, "fill");
  }
  if (V > 0) {
    hipLaunchKernelGGL(k_ad_normals, dim3(ad_blocks(V)), dim3(256), 0, ctx->stream, verts.as<double>(),
                       tris.as<uint32_t>(), voff.as<int>(), deg.as<int>(), ent.as<int>(), (int)V, rm->d_n);
    hipLaunchKernelGGL(k_ad_curvature, dim3(ad_blocks(V)), dim3(256), 0, ctx->stream, gv, p, verts.as<double>(), (int)V,
                       rm->d_c, 0);
  }
  if ((e = hipGetLastError()) != hipSuccess ||
      (V && (e = hipMemcpyAsync(rm->d_v, verts.p, sizeof(double) * 3 * V, hipMemcpyDeviceToDevice, ctx->stream)) !=
                hipSuccess) ||
      (T && (e = hipMemcpyAsync(rm->d_t, tris.p, sizeof(uint32_t) * 3 * T, hipMemcpyDeviceToDevice, ctx->stream)) !=
                hipSuccess) ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return vr_hip_fail(ctx, e, "adaptive surface");
  rm->nv = (uint64_t)V;
  rm->nt = (uint64_t)T;
  *out = rm.release();
  return VR_OK;
}

extern "C" int vr_render_mesh_counts(vr_ctx* ctx, const vr_render_mesh* m, uint64_t* vertex_count,
                                     uint64_t* tri_count) {
  if (!m) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null mesh");
  if (vertex_count) *vertex_count = m->nv;
  if (tri_count) *tri_count = m->nt;
  return VR_OK;
}

extern "C" int vr_render_mesh_copy(vr_ctx* ctx, const vr_render_mesh* m, double* xyz, double* normals,
                                   double* curvatures, size_t vcap, uint32_t* tri, size_t tcap) {
  if (!ctx || !m) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null argument");
  if (((xyz || normals || curvatures) && vcap < m->nv) || (tri && tcap < m->nt))
    return vr_fail(ctx, VR_ERR_CAPACITY, "mesh needs %llu vertices, %llu triangles", (unsigned long long)m->nv,
                   (unsigned long long)m->nt);
  hipError_t e;
  const size_t nv = (size_t)m->nv, nt = (size_t)m->nt;
  if ((xyz && nv && (e = hipMemcpyAsync(xyz, m->d_v, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, ctx->stream)) !=
                        hipSuccess) ||
      (normals && nv &&
       (e = hipMemcpyAsync(normals, m->d_n, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) ||
      (curvatures && nv &&
       (e = hipMemcpyAsync(curvatures, m->d_c, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) ||
      (tri && nt && (e = hipMemcpyAsync(tri, m->d_t, sizeof(uint32_t) * 3 * nt, hipMemcpyDeviceToHost, ctx->stream)) !=
                        hipSuccess))
    return vr_hip_fail(ctx, e, "hipMemcpyAsync");
  return (e = hipStreamSynchronize(ctx->stream)) == hipSuccess ? VR_OK : vr_hip_fail(ctx, e, "sync");
}

extern "C" int vr_render_mesh_destroy(vr_ctx* ctx, vr_render_mesh* m) {
  if (!m) return VR_ERR_INVALID_ARGUMENT;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete m;
  return VR_OK;
}
