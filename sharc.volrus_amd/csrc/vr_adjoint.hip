// Differentiable rollout: forward advection steps through a static dense
// velocity field with checkpoints, then the adjoint (reverse) pass.
//
// No reference counterpart (the reference has no differentiation).  The
// step is phi_{n+1}(x_i) = S(phi_n, p_i(U)), S the trilinear sample
// (src/tools/interpolation.rs:11-48) and p_i the RK1 or RK3 departure point
// through the dense field U.  Only active voxels are variables; inactive and
// background values are constants.  Per backward step and per leaf (one
// block of 4 waves):
//   - lambda_{n+1} of each active voxel is the sum of the brick entries
//     (leaf + one-voxel halo) that the previous backward step left for it
//     in the (up to 8) leaves whose brick covers it, in a fixed order;
//   - it is scattered with the trilinear weights into this leaf's brick as
//     exact 32-bit fixed-point sums, so lambda_n is deterministic;
//   - g_n = lambda_{n+1} dS/dp is added to the voxel's running sum over the
//     steps.
// The departure point depends on the voxel, the static field and dt only,
// so dL/dU = sum over voxels of (dp/dU)^T (sum over steps of g_n): after the
// last backward step one pass pulls each voxel's sum back through the
// departure point (and, for RK3, through the stage positions with the
// field's Jacobian) onto the velocity nodes, reduced in an LDS window
// (separably for the voxel's own position) and added to dL/dU with global
// atomics.
// Forward states are kept every `checkpoint_every` steps; the others are
// recomputed segment by segment during the reverse pass.
#include <cstring>
#include <vector>

#include "vr_advect.h"
#include "vr_internal.h"

#define VW 6  // velocity nodes per axis in a leaf's LDS window
#define VW3 (VW * VW * VW)

struct AdjStep {
  GridView phi;  // phi_n
  VelView vel;   // dense, static
  const float* ob_in;  // [cap][1000] bricks of step n+1 (lambda_{n+1} = their merge)
  float* far_in;       // [cap][512] far-corner sums of step n+1 (zeroed once read)
  float* ob_out;       // [cap][1000] bricks of this step
  float* far_out;      // [cap][512], corners beyond the brick (large steps)
  float* grad;         // 3 n^3
  // Per active voxel (leaf li's i-th active voxel in offset order at
  // abase[li] + i): the sum over the steps of lambda_{n+1} grad S(phi_n, p),
  // p the voxel's departure point; component a at gacc + a * gstride.
  const int* abase;
  float* gacc;
  long long gstride;
  double dt, tf0, tf1, tf2;
  int separable;   // grid transform diagonal with positive scales (host check)
};

// Trilinear node weights of the dense field at p, in double, and the base node.
__device__ __forceinline__ void dense_weights(const VelView& v, const double p[3], int b[3], double f[3]) {
  for (int a = 0; a < 3; ++a) {
    const double idx = p[a] * v.inv_h;
    const double fl = floor(idx);
    b[a] = f64_to_i32(fl);
    f[a] = idx - fl;
  }
}

// LDS float add on a pointer into __shared__ memory, always as a DS
// instruction (a generic-pointer atomic that may also target global memory
// would compile to a slower FLAT atomic).
__device__ __forceinline__ void lds_add(float* p, float v) {
  __builtin_amdgcn_ds_faddf((__attribute__((address_space(3))) float*)p, v, 0, 0, false);
}

// Adds coef[c] * w_q(p) to node q of component c: LDS window when inside it,
// else global.  The branch is taken once per point (all eight nodes inside
// the window, the usual case, or not), so the LDS path carries no global
// accesses.  Lanes visit the 8 nodes in a lane-rotated order: voxels of one
// wave mostly share a velocity cell, and the rotation makes the lanes of one
// atomic instruction hit different nodes instead of serialising on one.
__device__ __forceinline__ void vel_scatter(const AdjStep& P, float* vwin, const int wb[3], const double p[3],
                                            const double coef[3]) {
  int b[3];
  double f[3];
  dense_weights(P.vel, p, b, f);
  const int n = P.vel.n;
  const int lane = threadIdx.x & 63;
  const int rot = lane + (lane >> 3);
  const int wi = b[0] - wb[0], wj = b[1] - wb[1], wk = b[2] - wb[2];
  // Fast path: all eight nodes inside both the window and the field (nodes
  // outside the field carry no variable and get nothing).
  if ((unsigned)wi < VW - 1 && (unsigned)wj < VW - 1 && (unsigned)wk < VW - 1 && (unsigned)b[0] < (unsigned)(n - 1) &&
      (unsigned)b[1] < (unsigned)(n - 1) && (unsigned)b[2] < (unsigned)(n - 1)) {
    const int w0 = (wi * VW + wj) * VW + wk;
    for (int s = 0; s < 8; ++s) {
      const int q = (s + rot) & 7;
      const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
      const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
      const int o = w0 + (dx * VW + dy) * VW + dz;
      lds_add(&vwin[o], (float)(coef[0] * w));
      lds_add(&vwin[VW3 + o], (float)(coef[1] * w));
      lds_add(&vwin[2 * VW3 + o], (float)(coef[2] * w));
    }
    return;
  }
  const size_t n3 = (size_t)n * n * n;
  for (int q = 0; q < 8; ++q) {
    const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
    const int i = b[0] + dx, j = b[1] + dy, k = b[2] + dz;
    if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n || (unsigned)k >= (unsigned)n) continue;
    const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
    for (int c = 0; c < 3; ++c) atomicAdd(&P.grad[c * n3 + ((size_t)i * n + j) * n + k], (float)(coef[c] * w));
  }
}

// J^T g for the dense field at p: (J^T g)_a = sum_c g_c du_c/dp_a.
__device__ __forceinline__ void dense_jt(const VelView& v, const double p[3], const double g[3], double out[3]) {
  int b[3];
  double f[3];
  dense_weights(v, p, b, f);
  const size_t n3 = (size_t)v.n * v.n * v.n;
  out[0] = out[1] = out[2] = 0.0;
  for (int q = 0; q < 8; ++q) {
    const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
    const double wx = dx ? f[0] : 1.0 - f[0], wy = dy ? f[1] : 1.0 - f[1], wz = dz ? f[2] : 1.0 - f[2];
    const double sx = dx ? 1.0 : -1.0, sy = dy ? 1.0 : -1.0, sz = dz ? 1.0 : -1.0;
    double gu = 0.0;
    for (int c = 0; c < 3; ++c)
      gu += g[c] * (double)dense_at(v.dense + c * n3, v.n, b[0] + dx, b[1] + dy, b[2] + dz);
    out[0] += gu * sx * wy * wz;
    out[1] += gu * wx * sy * wz;
    out[2] += gu * wx * wy * sz;
  }
  for (int a = 0; a < 3; ++a) out[a] *= v.inv_h;
}

// The dense field at p (vel_sample<VK_DENSE>) with the cell's nodes read from
// s_u, the leaf's window of nodes wb + (0..VW-1)^3 (0 outside the field, as
// dense_at), when the cell lies in it: the same values, so the same bits.
__device__ __forceinline__ void win_vel(const VelView& v, const float* s_u, const int wb[3], const double p[3],
                                        double out[3]) {
  const double ix = p[0] * v.inv_h, iy = p[1] * v.inv_h, iz = p[2] * v.inv_h;
  const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
  const float fx = (float)(ix - flx), fy = (float)(iy - fly), fz = (float)(iz - flz);
  const int ri = f64_to_i32(flx) - wb[0], rj = f64_to_i32(fly) - wb[1], rk = f64_to_i32(flz) - wb[2];
  if ((unsigned)ri < VW - 1 && (unsigned)rj < VW - 1 && (unsigned)rk < VW - 1) {
    const int o = (ri * VW + rj) * VW + rk;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      float c[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) c[q] = s_u[comp * VW3 + o + ((q & 1) * VW + ((q >> 1) & 1)) * VW + ((q >> 2) & 1)];
      out[comp] = (double)trilinear_blend(c, fx, fy, fz);
    }
  } else {
    vel_sample<VK_DENSE>(v, p, out);
  }
}

// dense_jt with the nodes from the window when the cell lies in it.
__device__ __forceinline__ void win_jt(const VelView& v, const float* s_u, const int wb[3], const double p[3],
                                       const double g[3], double out[3]) {
  int b[3];
  double f[3];
  dense_weights(v, p, b, f);
  const int ri = b[0] - wb[0], rj = b[1] - wb[1], rk = b[2] - wb[2];
  if (!((unsigned)ri < VW - 1 && (unsigned)rj < VW - 1 && (unsigned)rk < VW - 1)) {
    dense_jt(v, p, g, out);
    return;
  }
  const int o = (ri * VW + rj) * VW + rk;
  out[0] = out[1] = out[2] = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
    const double wx = dx ? f[0] : 1.0 - f[0], wy = dy ? f[1] : 1.0 - f[1], wz = dz ? f[2] : 1.0 - f[2];
    const double sx = dx ? 1.0 : -1.0, sy = dy ? 1.0 : -1.0, sz = dz ? 1.0 : -1.0;
    const int oq = o + (dx * VW + dy) * VW + dz;
    double gu = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) gu += g[c] * (double)s_u[c * VW3 + oq];
    out[0] += gu * sx * wy * wz;
    out[1] += gu * wx * sy * wz;
    out[2] += gu * wx * wy * sz;
  }
  for (int a = 0; a < 3; ++a) out[a] *= v.inv_h;
}

#define ADJ_WAVES 4
#define ADJ_THREADS (64 * ADJ_WAVES)

// LDS 64-bit integer add (DS instruction).
__device__ __forceinline__ void lds_add_i64(long long* p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Trilinear weight of node `node` along one axis for a point with base node b
// and fraction f: 1 - f at b, f at b + 1, 0 elsewhere.
__device__ __forceinline__ double axis_weight(int node, int b, double f) {
  return node == b ? 1.0 - f : (node == b + 1 ? f : 0.0);
}

// lambda(y) for voxel (x, y, z) of leaf li: the brick entries of the leaves
// whose brick covers it, summed in neighbour order, plus the far-corner sum.
// A neighbour at offset (dx, dy, dz) covers the voxel only where its local
// coordinate is 0 (for d = -1) or 7 (for d = +1) on every axis with d != 0,
// so at most 8 bricks are visited, in ascending neighbour order.
// The (up to 8) candidates are visited as a fixed 2 x 2 x 2 walk whose
// entry (i, j, k) takes the i-th offset of the axis's ascending list ([0],
// [-1, 0] at local 0, [0, 1] at local 7): all loads are issued before the
// ordered sum, and an absent or unused candidate adds +0 (the sum starts at
// +0 and is never -0, so adding +0 leaves it unchanged).
__device__ __forceinline__ float merged_lambda(const float* ob, const float* far, const int* nb, int li, int x, int y,
                                               int z) {
  const int ox0 = x == 0 ? -1 : 0, oy0 = y == 0 ? -1 : 0, oz0 = z == 0 ? -1 : 0;
  const int nx = (x == 0 || x == 7) ? 2 : 1, ny = (y == 0 || y == 7) ? 2 : 1, nz = (z == 0 || z == 7) ? 2 : 1;
  const float fv = far[(size_t)li * 512 + ((x << 6) | (y << 3) | z)];
  float v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = q >> 2, j = (q >> 1) & 1, k = q & 1;
    const int dx = ox0 + i, dy = oy0 + j, dz = oz0 + k;
    const bool use = i < nx && j < ny && k < nz;
    const int src = use ? nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] : -1;
    const int lx = x - 8 * dx, ly = y - 8 * dy, lz = z - 8 * dz;
    // an unused candidate reads the own brick's entry, which is discarded
    const size_t at = src >= 0 ? (size_t)src * 1000 + (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1)
                               : (size_t)li * 1000 + (x + 1) * 100 + (y + 1) * 10 + (z + 1);
    const float t = ob[at];
    v[q] = src >= 0 ? t : 0.0f;
  }
  float s = 0.0f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s += v[q];
  return s + fv;
}

// The leaf's active voxels in offset order (list[0..total)), by wave.
__device__ __forceinline__ int adj_active_list(const NanoLeaf& L, uint16_t* list, int lane, int wave) {
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int total = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const uint64_t m = L.mask[x];
    if ((x % ADJ_WAVES) == wave && ((m >> lane) & 1ull)) list[total + __popcll(m & below)] = (uint16_t)(x * 64 + lane);
    total += __popcll(m);
  }
  return total;
}

// The dense field's window of the leaf at (ox, oy, oz): base node wb (two
// nodes below the leaf origin's cell) and its nodes wb + (0..VW-1)^3 in s_u
// (0 outside the field, as dense_at).  A thread's nodes all load before any
// is stored (nodes outside the field read node 0).
__device__ __forceinline__ void adj_window(const GridView& g, const VelView& v, int ox, int oy, int oz, int wb[3],
                                           float* s_u, int tid) {
  double ow[3];
  index_to_world(g.mat, ox, oy, oz, ow);
  for (int a = 0; a < 3; ++a) wb[a] = f64_to_i32(floor(ow[a] * v.inv_h)) - 2;
  constexpr int UQ = (3 * VW3 + ADJ_THREADS - 1) / ADJ_THREADS;
  const int nn = v.n;
  const size_t n3 = (size_t)nn * nn * nn;
  float uv[UQ];
  bool inside[UQ];
#pragma unroll
  for (int q = 0; q < UQ; ++q) {
    const int i0 = tid + q * ADJ_THREADS, i = i0 < 3 * VW3 ? i0 : 0;
    const int c = i / VW3, r = i - c * VW3;
    const int gi = wb[0] + r / (VW * VW), gj = wb[1] + (r / VW) % VW, gk = wb[2] + r % VW;
    inside[q] = (unsigned)gi < (unsigned)nn && (unsigned)gj < (unsigned)nn && (unsigned)gk < (unsigned)nn;
    uv[q] = v.dense[inside[q] ? c * n3 + ((size_t)gi * nn + gj) * nn + gk : 0];
  }
#pragma unroll
  for (int q = 0; q < UQ; ++q) {
    const int i0 = tid + q * ADJ_THREADS;
    if (i0 < 3 * VW3) s_u[i0] = inside[q] ? uv[q] : 0.0f;
  }
}

// The departure point of voxel (x, y, z) of the leaf at (ox, oy, oz): RK1 or
// RK3 through the static field (the forward step's arithmetic, so the same
// point), with the RK3 stage positions y1, y2.
template <int SCHEME>
__device__ __forceinline__ void adj_departure(const AdjStep& P, const float* s_u, const int wb[3], int px, int py,
                                              int pz, const double tf[3], double xw[3], double y1[3], double y2[3],
                                              double p[3]) {
  double u[3];
  index_to_world(P.phi.mat, px, py, pz, xw);
  win_vel(P.vel, s_u, wb, xw, u);
  for (int a = 0; a < 3; ++a) y1[a] = xw[a] - P.dt * (u[a] * tf[0]);
  if (SCHEME == 1) {
    for (int a = 0; a < 3; ++a) p[a] = y1[a];
  } else {
    win_vel(P.vel, s_u, wb, y1, u);
    for (int a = 0; a < 3; ++a) y2[a] = 0.75 * xw[a] + 0.25 * (y1[a] - P.dt * (u[a] * tf[1]));
    win_vel(P.vel, s_u, wb, y2, u);
    for (int a = 0; a < 3; ++a) p[a] = (1.0 / 3.0) * xw[a] + (2.0 / 3.0) * (y2[a] - P.dt * (u[a] * tf[2]));
  }
}

// One backward step for the state adjoint, one leaf per block of ADJ_WAVES
// waves.
//
// lambda_{n+1} of the leaf's voxels is first merged from the previous step's
// bricks (so no separate merge launch), then scattered with the trilinear
// weights of the departure cell into this leaf's brick: each contribution
// (float)(lambda w) is quantised to an integer with the leaf's scale
// 2^(21 - E), max |lambda_{n+1}| over the leaf < 2^E, and summed in 32 bits
// (a brick cell gets at most one per voxel, 512, so the sum stays below
// 2^30).  Integer sums are exact, so lambda_n does not depend on the order of
// the adds, and integer LDS atomics run far faster than float ones.
//
// The velocity gradient of the step is (dp/dU)^T g_n per voxel, with
// g_n = lambda_{n+1} grad S(phi_n, p) and dp/dU the departure point's
// derivative, which depends on the voxel, the static field and dt only: the
// same every step.  So each voxel's g_n is summed over the steps (gacc) and
// pulled back through dp/dU once, after the last backward step
// (adj_vel_step), instead of once per step.
template <int SCHEME>
__device__ __forceinline__ void adj_lambda_step(const AdjStep& P, int li, bool valid) {
  __shared__ float brick[1000];
  __shared__ int acc[1000];
  __shared__ float s_u[3 * VW3];  // the field's nodes over the window
  __shared__ uint16_t list[512];
  __shared__ int nb[VR_NBR];
  __shared__ float s_max[ADJ_WAVES];
  __shared__ float s_lam[512];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const GridView& g = P.phi;
  const double tf[3] = {P.tf0, P.tf1, P.tf2};
  if (valid) {
    const NanoLeaf& L = g.leaves[li];
    if (tid < VR_NBR) nb[tid] = g.nbr[(size_t)li * VR_NBR + tid];
    const int total = adj_active_list(L, list, lane, wave);
    for (int i = tid; i < 1000; i += ADJ_THREADS) acc[i] = 0;
    const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
    int wb[3];
    adj_window(g, P.vel, ox, oy, oz, wb, s_u, tid);
    stage_interior<ADJ_THREADS>(g, li, brick, tid);
    __syncthreads();
    // lambda_{n+1} of the active voxels (merged), its max |.|; the leaf's
    // far-corner entries are consumed and cleared for the next step.
    float lm = 0.0f;
    for (int off = tid; off < 512; off += ADJ_THREADS) {
      float lam = 0.0f;
      if ((L.mask[off >> 6] >> (off & 63)) & 1ull)
        lam = merged_lambda(P.ob_in, P.far_in, nb, li, off >> 6, (off >> 3) & 7, off & 7);
      s_lam[off] = lam;
      P.far_in[(size_t)li * 512 + off] = 0.0f;
      lm = fmaxf(lm, fabsf(lam));
    }
    for (int o = 32; o > 0; o >>= 1) lm = fmaxf(lm, __shfl_xor(lm, o));
    if (lane == 0) s_max[wave] = lm;
    stage_halo<ADJ_THREADS>(g, nb, brick, tid);
    __syncthreads();
    float mx = s_max[0];
#pragma unroll
    for (int w = 1; w < ADJ_WAVES; ++w) mx = fmaxf(mx, s_max[w]);
    int ex = 0;
    (void)frexpf(mx, &ex);  // mx < 2^ex
    // contributions |v| < 2^ex quantise to |q| < 2^21; a cell gets at most
    // one per voxel (a voxel's 8 corners are distinct), so at most 512 sum
    // below 2^30 in int32
    const float scale = ldexpf(1.0f, 21 - ex);
    const double inv_scale = ldexp(1.0, ex - 21);
    float* const ga = P.gacc + P.abase[li];
    for (int ii = tid; ii < total && mx > 0.0f; ii += ADJ_THREADS) {
      const int off = list[ii];
      const float lam = s_lam[off];
      if (lam == 0.0f) continue;
      const int x = off >> 6, y = (off >> 3) & 7, z = off & 7;
      double xw[3], y1[3], y2[3], p[3];
      adj_departure<SCHEME>(P, s_u, wb, ox + x, oy + y, oz + z, tf, xw, y1, y2, p);
      const double ix = p[0] * g.inv_vs, iy = p[1] * g.inv_vs, iz = p[2] * g.inv_vs;
      const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
      const double fx = ix - flx, fy = iy - fly, fz = iz - flz;
      const int i0 = f64_to_i32(flx), j0 = f64_to_i32(fly), k0 = f64_to_i32(flz);
      const int bi = i0 - ox, bj = j0 - oy, bk = k0 - oz;
      const bool inb = bi >= -1 && bi <= 7 && bj >= -1 && bj <= 7 && bk >= -1 && bk <= 7;
      double c[8];
      if (inb) {
        // Departure cell inside the LDS brick (the usual case).
        const int b0 = (bi + 1) * 100 + (bj + 1) * 10 + (bk + 1);
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = (double)brick[b0 + (q & 1) * 100 + ((q >> 1) & 1) * 10 + ((q >> 2) & 1)];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
          const double w = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
          const float v = (float)((double)lam * w);
          atomicAdd(&acc[b0 + dx * 100 + dy * 10 + dz], (int)rintf(v * scale));
        }
      } else {
        for (int q = 0; q < 8; ++q) {
          const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
          const double w = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
          const int cx = i0 + dx, cy = j0 + dy, cz = k0 + dz;
          const int tl = find_leaf(g, cx, cy, cz);
          c[q] = tl < 0 ? (double)g.background : (double)g.leaves[tl].values[voxel_offset(cx, cy, cz)];
          if (tl >= 0) atomicAdd(&P.far_out[(size_t)tl * 512 + voxel_offset(cx, cy, cz)], (float)((double)lam * w));
        }
      }
      const double dfx = (c[1] - c[0]) * (1 - fy) * (1 - fz) + (c[3] - c[2]) * fy * (1 - fz) +
                         (c[5] - c[4]) * (1 - fy) * fz + (c[7] - c[6]) * fy * fz;
      const double dfy = (c[2] - c[0]) * (1 - fx) * (1 - fz) + (c[3] - c[1]) * fx * (1 - fz) +
                         (c[6] - c[4]) * (1 - fx) * fz + (c[7] - c[5]) * fx * fz;
      const double dfz = (c[4] - c[0]) * (1 - fx) * (1 - fy) + (c[5] - c[1]) * fx * (1 - fy) +
                         (c[6] - c[2]) * (1 - fx) * fy + (c[7] - c[3]) * fx * fy;
      // this step's g_n, summed into the voxel's gacc entry (one thread per
      // voxel and step, so the sum is in step order)
      ga[ii] += (float)((double)lam * dfx * g.inv_vs);
      ga[P.gstride + ii] += (float)((double)lam * dfy * g.inv_vs);
      ga[2 * P.gstride + ii] += (float)((double)lam * dfz * g.inv_vs);
    }
    __syncthreads();
    float* ob = P.ob_out + (size_t)li * 1000;
    for (int i = tid; i < 1000; i += ADJ_THREADS) ob[i] = (float)((double)acc[i] * inv_scale);
    __syncthreads();
  }
}

// The velocity gradient of the whole rollout, one leaf per block: each
// active voxel's summed g (gacc) pulled back through its departure point's
// derivative dp/dU.
//
// Scatter at the voxel's own position (the RK1 term and the last RK3 term):
// for a diagonal transform the node weights factor per axis, so the
// per-voxel coefficients are staged in LDS and contracted axis by axis into
// the leaf's node window (no atomics); otherwise float atomics into the
// window.  The two RK3 stage positions are kept as records and added in a
// second pass as int32-quantised 64-bit sums scaled by the leaf's max
// |coefficient|.  The window is then added to dL/dU with global atomics.
template <int SCHEME>
__device__ __forceinline__ void adj_vel_step(const AdjStep& P, int li, bool valid) {
  // RK3 keeps 2 scatter records per voxel (y2, y1); the scatter at x goes
  // through the separable contraction, or straight into the window.
  constexpr int NREC = SCHEME == 3 ? 2 : 0;
  constexpr int REC_WORDS = NREC * 6 * 512;
  constexpr int T_WORDS = 3 * 64 * VW + 3 * 8 * VW * VW;
  __shared__ float coefx[3][512];
  __shared__ float vwin[3 * VW3];
  __shared__ float s_u[3 * VW3];  // the field's nodes over the window
  __shared__ long long vwin_i[SCHEME == 3 ? 3 * VW3 : 1];
  // records (pass 1 -> pass 2), then the separable contraction stages
  __shared__ float s_raw[REC_WORDS > T_WORDS ? REC_WORDS : T_WORDS];
  float* const t1 = s_raw;
  float* const t2 = s_raw + 3 * 64 * VW;
  // record r of a voxel: per axis the coefficient, and the cell's base node
  // relative to the window (10 bits) with the fraction in 22-bit fixed point
  // above it; all-zero coefficients mark no record.
  float* const rec_c = s_raw;                                  // [NREC][3][512]
  uint32_t* const rec_w = (uint32_t*)(s_raw + NREC * 3 * 512);  // [NREC][3][512]
  __shared__ float s_cmax[ADJ_WAVES];
  __shared__ uint16_t list[512];
  // per axis and local coordinate: base node relative to the window, and the
  // weight of each window node
  __shared__ int s_bn[3][8];
  __shared__ float s_wt[3][8][VW];
  __shared__ int s_sep;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const GridView& g = P.phi;
  const double tf[3] = {P.tf0, P.tf1, P.tf2};
  if (valid) {
    const NanoLeaf& L = g.leaves[li];
    const int total = adj_active_list(L, list, lane, wave);
    for (int i = tid; i < 3 * 512; i += ADJ_THREADS) (&coefx[0][0])[i] = 0.0f;
    for (int i = tid; i < 3 * VW3; i += ADJ_THREADS) vwin[i] = 0.0f;
    if constexpr (SCHEME == 3) {
      for (int i = tid; i < 3 * VW3; i += ADJ_THREADS) vwin_i[i] = 0;
      for (int i = tid; i < NREC * 3 * 512; i += ADJ_THREADS) rec_c[i] = 0.0f;
    }
    const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
    int wb[3];
    adj_window(g, P.vel, ox, oy, oz, wb, s_u, tid);
    if (tid < 24) {
      // Node base and fraction per axis of the voxel positions (diagonal
      // transform: the axis coordinate alone fixes them).
      const int a = tid >> 3, l = tid & 7;
      const int o = a == 0 ? ox : (a == 1 ? oy : oz);
      const double w = g.mat[5 * a] * (double)(o + l) + g.mat[4 * a + 3];
      const double idx = w * P.vel.inv_h;
      const double fl = floor(idx);
      const int bn = f64_to_i32(fl) - wb[a];
      s_bn[a][l] = bn;
      for (int k = 0; k < VW; ++k) s_wt[a][l][k] = (float)axis_weight(k, bn, idx - fl);
    }
    __syncthreads();
    if (tid == 0) {
      bool ok = P.separable != 0;
      for (int a = 0; a < 3; ++a) ok = ok && s_bn[a][0] >= 0 && s_bn[a][7] + 1 < VW;
      s_sep = ok ? 1 : 0;
    }
    __syncthreads();
    const bool sep = s_sep != 0;
    float cmax = 0.0f;  // max |coefficient| over this thread's records
    // Stores a scatter record (RK3), or adds far points straight to dL/dU.
    auto rec_put = [&](int r, int off, const double pt[3], const double c[3]) {
      int b[3];
      double f[3];
      dense_weights(P.vel, pt, b, f);
      const int r0 = b[0] - wb[0], r1 = b[1] - wb[1], r2 = b[2] - wb[2];
      if (r0 < -512 || r0 > 511 || r1 < -512 || r1 > 511 || r2 < -512 || r2 > 511) {
        const int nn = P.vel.n;
        const size_t n3 = (size_t)nn * nn * nn;
        for (int q = 0; q < 8; ++q) {
          const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
          const int i = b[0] + dx, j = b[1] + dy, k = b[2] + dz;
          if ((unsigned)i >= (unsigned)nn || (unsigned)j >= (unsigned)nn || (unsigned)k >= (unsigned)nn) continue;
          const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
          for (int a = 0; a < 3; ++a) atomicAdd(&P.grad[a * n3 + ((size_t)i * nn + j) * nn + k], (float)(c[a] * w));
        }
        return;
      }
      const int rr[3] = {r0, r1, r2};
      for (int a = 0; a < 3; ++a) {
        // f in [0, 1): 22-bit fixed point, truncated
        const uint32_t fq = (uint32_t)(f[a] * 4194304.0);
        rec_w[(r * 3 + a) * 512 + off] = (fq << 10) | (uint32_t)(rr[a] & 1023);
        const float cf = (float)c[a];
        rec_c[(r * 3 + a) * 512 + off] = cf;
        cmax = fmaxf(cmax, fabsf(cf));
      }
    };
    const float* const ga = P.gacc + P.abase[li];
    for (int ii = tid; ii < total; ii += ADJ_THREADS) {
      const int off = list[ii];
      const double gp[3] = {(double)ga[ii], (double)ga[P.gstride + ii], (double)ga[2 * P.gstride + ii]};
      if (gp[0] == 0.0 && gp[1] == 0.0 && gp[2] == 0.0) continue;
      const int x = off >> 6, y = (off >> 3) & 7, z = off & 7;
      double xw[3], y1[3], y2[3], p[3];
      adj_departure<SCHEME>(P, s_u, wb, ox + x, oy + y, oz + z, tf, xw, y1, y2, p);
      double cself[3];  // coefficient of the scatter at the voxel's own position
      if (SCHEME == 1) {
        for (int a = 0; a < 3; ++a) cself[a] = -P.dt * tf[0] * gp[a];
      } else {
        double jt[3], gy2[3], gy1[3];
        win_jt(P.vel, s_u, wb, y2, gp, jt);
        for (int a = 0; a < 3; ++a) gy2[a] = (2.0 / 3.0) * gp[a] - (2.0 / 3.0) * P.dt * tf[2] * jt[a];
        const double c3[3] = {-(2.0 / 3.0) * P.dt * tf[2] * gp[0], -(2.0 / 3.0) * P.dt * tf[2] * gp[1],
                              -(2.0 / 3.0) * P.dt * tf[2] * gp[2]};
        rec_put(0, off, y2, c3);
        win_jt(P.vel, s_u, wb, y1, gy2, jt);
        for (int a = 0; a < 3; ++a) gy1[a] = 0.25 * gy2[a] - 0.25 * P.dt * tf[1] * jt[a];
        const double c2[3] = {-0.25 * P.dt * tf[1] * gy2[0], -0.25 * P.dt * tf[1] * gy2[1],
                              -0.25 * P.dt * tf[1] * gy2[2]};
        rec_put(1, off, y1, c2);
        for (int a = 0; a < 3; ++a) cself[a] = -P.dt * tf[0] * gy1[a];
      }
      if (sep) {
        for (int a = 0; a < 3; ++a) coefx[a][off] = (float)cself[a];
      } else {
        vel_scatter(P, vwin, wb, xw, cself);
      }
    }
    double inv_scale_v = 0.0;
    if constexpr (SCHEME == 3) {
      // Pass 2: records into the node window as exact 64-bit fixed-point
      // sums of int32-quantised contributions.
      for (int o = 32; o > 0; o >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, o));
      if (lane == 0) s_cmax[wave] = cmax;
      __syncthreads();
      float cm = s_cmax[0];
#pragma unroll
      for (int w = 1; w < ADJ_WAVES; ++w) cm = fmaxf(cm, s_cmax[w]);
      int ev = 0;
      (void)frexpf(cm, &ev);
      // |c w| < 2^ev quantises to int32 (|q| < 2^30); at most 3 x 512
      // contributions per node sum below 2^41 in int64
      const float scale_v = ldexpf(1.0f, 30 - ev);
      inv_scale_v = ldexp(1.0, ev - 30);
      const int nn = P.vel.n;
      const size_t n3 = (size_t)nn * nn * nn;
      const int rot = lane + (lane >> 3);
      for (int ii = tid; ii < total; ii += ADJ_THREADS) {
        const int off = list[ii];
        for (int r = 0; r < NREC; ++r) {
          float f[3], cs[3];
          int rel[3];
          for (int a = 0; a < 3; ++a) {
            const uint32_t wv = rec_w[(r * 3 + a) * 512 + off];
            rel[a] = ((int)(wv << 22)) >> 22;
            f[a] = (float)(wv >> 10) * 2.384185791015625e-07f;  // 2^-22
            cs[a] = rec_c[(r * 3 + a) * 512 + off] * scale_v;
          }
          if (cs[0] == 0.0f && cs[1] == 0.0f && cs[2] == 0.0f) continue;
          const bool inwin = (unsigned)rel[0] < VW - 1 && (unsigned)rel[1] < VW - 1 && (unsigned)rel[2] < VW - 1 &&
                             (unsigned)(wb[0] + rel[0]) < (unsigned)(nn - 1) &&
                             (unsigned)(wb[1] + rel[1]) < (unsigned)(nn - 1) &&
                             (unsigned)(wb[2] + rel[2]) < (unsigned)(nn - 1);
          if (inwin) {
            // all 8 nodes in the window and the field: nodes and components
            // in a lane-rotated order, so lanes sharing a cell spread over
            // 24 addresses
            const int o0 = (rel[0] * VW + rel[1]) * VW + rel[2];
            const int crot = (lane >> 3) % 3;
#pragma unroll 1
            for (int st = 0; st < 8; ++st) {
              const int q = (st + rot) & 7;
              const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
              const float w = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
              const int o = o0 + (dx * VW + dy) * VW + dz;
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                const int cc = a + crot >= 3 ? a + crot - 3 : a + crot;
                const float c = cc == 0 ? cs[0] : (cc == 1 ? cs[1] : cs[2]);
                lds_add_i64(&vwin_i[cc * VW3 + o], (long long)(int)rintf(c * w));
              }
            }
          } else {
            for (int q = 0; q < 8; ++q) {
              const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
              const int wi = rel[0] + dx, wj = rel[1] + dy, wk = rel[2] + dz;
              const int gi = wb[0] + wi, gj = wb[1] + wj, gk = wb[2] + wk;
              if ((unsigned)gi >= (unsigned)nn || (unsigned)gj >= (unsigned)nn || (unsigned)gk >= (unsigned)nn) continue;
              const float w = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
              if ((unsigned)wi < VW && (unsigned)wj < VW && (unsigned)wk < VW) {
                const int o = (wi * VW + wj) * VW + wk;
                for (int a = 0; a < 3; ++a) lds_add_i64(&vwin_i[a * VW3 + o], (long long)(int)rintf(cs[a] * w));
              } else {
                for (int a = 0; a < 3; ++a)
                  atomicAdd(&P.grad[a * n3 + ((size_t)gi * nn + gj) * nn + gk], (float)((double)cs[a] * w * inv_scale_v));
              }
            }
          }
        }
      }
    }
    __syncthreads();
    if (sep) {
      // Separable contraction of coefx into the window: z, then y, then x,
      // over the nodes the leaf's voxels touch (per axis s_bn[a][0] ..
      // s_bn[a][7] + 1; two or three for fields coarser than the voxels).
      const int lx = s_bn[0][0], ly = s_bn[1][0], lz = s_bn[2][0];
      const int nx = s_bn[0][7] + 2 - lx, ny = s_bn[1][7] + 2 - ly, nz = s_bn[2][7] + 2 - lz;
      for (int i = tid; i < 3 * 64 * nz; i += ADJ_THREADS) {
        const int c = i / (64 * nz), r = i - c * 64 * nz, xy = r / nz, k = r - xy * nz;
        float sum = 0.0f;
#pragma unroll
        for (int zz = 0; zz < 8; ++zz) sum += coefx[c][xy * 8 + zz] * s_wt[2][zz][lz + k];
        t1[i] = sum;
      }
      __syncthreads();
      for (int i = tid; i < 3 * 8 * ny * nz; i += ADJ_THREADS) {
        const int c = i / (8 * ny * nz), r = i - c * 8 * ny * nz, xx = r / (ny * nz), jk = r - xx * ny * nz;
        const int j = jk / nz, k = jk - j * nz;
        float sum = 0.0f;
#pragma unroll
        for (int yy = 0; yy < 8; ++yy) sum += t1[(c * 64 + xx * 8 + yy) * nz + k] * s_wt[1][yy][ly + j];
        t2[i] = sum;
      }
      __syncthreads();
      for (int i = tid; i < 3 * nx * ny * nz; i += ADJ_THREADS) {
        const int c = i / (nx * ny * nz), r = i - c * nx * ny * nz, ii = r / (ny * nz), jk = r - ii * ny * nz;
        const int j = jk / nz, k = jk - j * nz;
        float sum = 0.0f;
#pragma unroll
        for (int xx = 0; xx < 8; ++xx) sum += t2[(c * 8 + xx) * ny * nz + jk] * s_wt[0][xx][lx + ii];
        vwin[c * VW3 + ((lx + ii) * VW + ly + j) * VW + lz + k] += sum;
      }
      __syncthreads();
    }
    const int nv = P.vel.n;
    const size_t n3 = (size_t)nv * nv * nv;
    for (int i = tid; i < 3 * VW3; i += ADJ_THREADS) {
      float v = vwin[i];
      if constexpr (SCHEME == 3) v += (float)((double)vwin_i[i] * inv_scale_v);
      if (v == 0.0f) continue;
      const int c = i / VW3, r = i % VW3;
      const int gi = wb[0] + r / (VW * VW), gj = wb[1] + (r / VW) % VW, gk = wb[2] + r % VW;
      if ((unsigned)gi >= (unsigned)nv || (unsigned)gj >= (unsigned)nv || (unsigned)gk >= (unsigned)nv) continue;
      atomicAdd(&P.grad[c * n3 + ((size_t)gi * nv + gj) * nv + gk], v);
    }
    __syncthreads();
  }
}

// One leaf per block: the grid has at least ceil(n / 8) blocks per XCD.
// PASS 0: the state adjoint step (adj_lambda_step), PASS 1: the velocity
// gradient (adj_vel_step).  The RK3 velocity pass keeps its registers (2
// waves per SIMD); the others are held to 128 VGPRs (4 waves).
template <int SCHEME, int PASS>
__device__ __forceinline__ void adj_run(const AdjStep& P, int li, bool valid) {
  if constexpr (PASS == 0)
    adj_lambda_step<SCHEME>(P, li, valid);
  else
    adj_vel_step<SCHEME>(P, li, valid);
}
#define ADJ_WPE(SCHEME, PASS) __attribute__((amdgpu_waves_per_eu((SCHEME == 3 && PASS == 1) ? 2 : 4, 8)))
template <int SCHEME, int PASS>
__global__ __launch_bounds__(ADJ_THREADS) ADJ_WPE(SCHEME, PASS) void k_adj(AdjStep P) {
  int li_end, li_step;
  const int li = xcd_first(*P.phi.count, &li_end, &li_step);
  adj_run<SCHEME, PASS>(P, li, li < li_end);
}

// Grid g of global leaf i: leaf_base[g] <= i < leaf_base[g + 1] (G + 1 entries).
__device__ __forceinline__ int adj_grid_of(const int* leaf_base, int G, int i) {
  int lo = 0, hi = G;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (leaf_base[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Several level sets in one launch: the leaves of all grids form one list
// (grid g owns [leaf_base[g], leaf_base[g + 1])), one leaf per block, each
// with its grid's step (its state, field, bricks and gradient).
template <int SCHEME, int PASS>
__global__ __launch_bounds__(ADJ_THREADS) ADJ_WPE(SCHEME, PASS) void k_adj_multi(const AdjStep* Ps,
                                                                                 const int* leaf_base, int G, int L) {
  int end, stride;
  const int gl = xcd_first(L, &end, &stride);
  const bool valid = gl < end;
  const int g = __builtin_amdgcn_readfirstlane(valid ? adj_grid_of(leaf_base, G, gl) : 0);
  adj_run<SCHEME, PASS>(Ps[g], gl - leaf_base[g], valid);
}

// Per-leaf active voxel counts (for the gacc offsets).
__global__ void k_adj_leaf_active(const NanoLeaf* leaves, int n, int* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int c = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) c += __popcll(leaves[i].mask[w]);
    out[i] = c;
  }
}

// The forward step of several level sets in one launch (the dense field's
// advection, one leaf per block of NW waves): grid g reads ins[g] and writes
// outs[g] with its own field and time step.
template <int SCHEME, int NW>
__global__ __launch_bounds__(64 * NW) void k_adj_advect_multi(const GridView* ins, NanoLeaf* const* outs,
                                                             const VelView* vels, const double* dts,
                                                             const int* leaf_base, int G, int L) {
  __shared__ AdvectLDS<VK_DENSE> s;
  int end, stride;
  const int gl = xcd_first(L, &end, &stride);
  if (gl >= end) return;
  const int g = adj_grid_of(leaf_base, G, gl);
  const GridView in = ins[g];
  const VelView v = vels[g];
  NanoLeaf* const out = (VR_GLOBAL NanoLeaf*)reinterpret_cast<const uintptr_t*>(outs)[g];
  const double tf[3] = {1.0, 1.0, 1.0};
  advect_leaf<VK_DENSE, SCHEME, NW>(in, v, out, gl - leaf_base[g], s, threadIdx.x, dts[g], tf);
}

// lambda_K = phi_K - target at active voxels (0 elsewhere), written as each
// leaf's brick (interior entries; the halo is 0) so the first backward step
// merges it like any other step's bricks; per-leaf loss in f64.  The target
// value of a voxel is Grid::get: the target leaf with the same origin holds
// all of the leaf's voxels, so it is looked up once per leaf.
__device__ __forceinline__ void adj_terminal_leaf(const GridView& g, const GridView& target, float* ob,
                                                  double* leaf_loss, int li, bool valid) {
  __shared__ double s_part[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (valid) {
    const NanoLeaf& L = g.leaves[li];
    const int tl = find_leaf(target, L.origin[0], L.origin[1], L.origin[2]);
    double part = 0.0;
    for (int i = tid; i < 1000; i += 256) {
      const int bx = i / 100 - 1, by = (i / 10) % 10 - 1, bz = i % 10 - 1;
      double d = 0.0;
      if (bx >= 0 && bx < 8 && by >= 0 && by < 8 && bz >= 0 && bz < 8 && ((L.mask[bx] >> (by * 8 + bz)) & 1ull)) {
        const int off = bx * 64 + by * 8 + bz;
        const float t = tl < 0 ? target.background : target.leaves[tl].values[off];
        d = (double)L.values[off] - (double)t;
        part += 0.5 * d * d;
      }
      ob[(size_t)li * 1000 + i] = (float)d;
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
    if (lane == 0) s_part[wave] = part;
    __syncthreads();
    if (tid == 0) leaf_loss[li] = ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
    __syncthreads();
  }
}

// One leaf per block: the grid has at least ceil(n / 8) blocks per XCD.
__global__ __launch_bounds__(256) void k_adj_terminal(GridView g, GridView target, float* ob, double* leaf_loss) {
  int li_end, li_step;
  const int li = xcd_first(*g.count, &li_end, &li_step);
  adj_terminal_leaf(g, target, ob, leaf_loss, li, li < li_end);
}

// Several level sets: grid g's bricks and per-leaf losses at its leaf_base offset.
__global__ __launch_bounds__(256) void k_adj_terminal_multi(const GridView* gs, const GridView* targets, float* ob,
                                                            double* leaf_loss, const int* leaf_base, int G, int L) {
  int end, stride;
  const int gl = xcd_first(L, &end, &stride);
  const bool valid = gl < end;
  const int g = valid ? adj_grid_of(leaf_base, G, gl) : 0;
  const GridView gv = gs[g], tv = targets[g];
  const size_t b = (size_t)leaf_base[g];
  adj_terminal_leaf(gv, tv, ob + b * 1000, leaf_loss + b, gl - leaf_base[g], valid);
}

// Sum of n doubles in a fixed order (one block).
__global__ __launch_bounds__(256) void k_sum_f64(const double* v, const int* count, double* out) {
  __shared__ double s[256];
  const int n = *count;
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a += v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_zero_f32(float* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

static int grid_blocks(size_t n, int b) {
  size_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// Blocks of the one-leaf-per-block step kernels: at least ceil(L / 8) on
// each XCD (the leaves are cut into 8 ranges).
static int leaf_blocks(size_t L) { return (int)(8 * ((L + 7) / 8 + 1)); }

extern "C" int vr_adjoint_rollout(vr_ctx* ctx, const vr_grid* phi0, const vr_velocity* vel, double dt, int scheme,
                                  int steps, int checkpoint_every, const vr_grid* target, vr_adjoint_stats* stats,
                                  float* grad_out) {
  if (!ctx || !phi0 || !vel || !target || steps < 1 || checkpoint_every < 1)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad adjoint arguments");
  if (vel->kind != VK_DENSE || vel->period > 0.0)
    return vr_fail(ctx, VR_ERR_UNSUPPORTED, "the adjoint needs a static dense velocity field");
  if (scheme != VR_SCHEME_RK1 && scheme != VR_SCHEME_RK3)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "scheme must be 1 or 3");
  if (phi0->topo->has_empty)
    return vr_fail(ctx, VR_ERR_UNSUPPORTED, "the adjoint needs leaves that all hold active voxels");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  const int L = phi0->topo->count;
  const int cap = L < 1 ? 1 : L;
  const int c = checkpoint_every;
  const int nseg = (steps + c - 1) / c;
  const size_t nv3 = (size_t)vel->n * vel->n * vel->n;
  // Buffers: checkpoints (segment starts 1..nseg-1), segment states, final state.
  std::vector<NanoLeaf*> ck((size_t)nseg, nullptr), seg((size_t)c, nullptr);
  NanoLeaf* fin = nullptr;
  float *ob[2] = {nullptr, nullptr}, *far[2] = {nullptr, nullptr}, *grad = nullptr, *gacc = nullptr;
  double *leaf_loss = nullptr, *loss_d = nullptr;
  int *abase = nullptr, *acnt = nullptr, *ascan = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  uint64_t lc0 = 0, ac0 = 0;
  int rc = vr_grid_counts(ctx, phi0, &lc0, &ac0);
  if (rc != VR_OK) return rc;
  if (ac0 > 0x7fffffffULL) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "too many active voxels");
  const size_t na = ac0 ? (size_t)ac0 : 1;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc != VR_OK) return;
    if ((e = hipMalloc(p, bytes)) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipMalloc");
  };
  alloc((void**)&gacc, sizeof(float) * 3 * na);
  alloc((void**)&abase, sizeof(int) * (cap + 1));
  alloc((void**)&acnt, sizeof(int) * cap);
  alloc((void**)&ascan, sizeof(int) * (vr_scan_scratch_ints(cap) + 1));
  ck[0] = phi0->d_leaves;
  for (int k = 1; k < nseg; ++k) alloc((void**)&ck[k], sizeof(NanoLeaf) * cap);
  for (int j = 1; j < c; ++j) alloc((void**)&seg[j], sizeof(NanoLeaf) * cap);
  alloc((void**)&fin, sizeof(NanoLeaf) * cap);
  for (int i = 0; i < 2; ++i) {
    alloc((void**)&ob[i], sizeof(float) * 1000 * cap);
    alloc((void**)&far[i], sizeof(float) * 512 * cap);
  }
  alloc((void**)&grad, sizeof(float) * 3 * nv3);
  alloc((void**)&leaf_loss, sizeof(double) * cap);
  alloc((void**)&loss_d, sizeof(double));
  for (int i = 0; i < 3 && rc == VR_OK; ++i)
    if ((e = hipEventCreate(&ev[i])) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipEventCreate");
  // A grid view over an arbitrary leaf buffer with phi0's topology.
  vr_grid view;
  copy_header_into(&view, phi0);
  view.topo = phi0->topo;
  const VelView vv = make_vel_view(vel);
  // Diagonal transform with positive scales: velocity-node weights of voxel
  // positions factor per axis.
  int separable = 1;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      if (r == c ? !(phi0->mat[r][c] > 0.0) : phi0->mat[r][c] != 0.0) separable = 0;
  const double tf[3] = {1.0, 1.0, 1.0};
  auto step_forward = [&](NanoLeaf* src, NanoLeaf* dst) {
    if (rc != VR_OK) return;
    view.d_leaves = src;
    if ((e = vr_launch_advect(make_view(&view), vv, dst, L, dt, tf, scheme, vr_advect_waves(L, ctx->cus),
                              ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "advect");
  };
  if (rc == VR_OK) {
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(3 * nv3, 256)), dim3(256), 0, ctx->stream, grad, 3 * nv3);
    for (int i = 0; i < 2; ++i)
      hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(512 * (size_t)cap, 256)), dim3(256), 0, ctx->stream, far[i],
                         512 * (size_t)cap);
    (void)hipEventRecord(ev[0], ctx->stream);
  }
  // Forward: states at segment starts are kept; the last state goes to `fin`.
  NanoLeaf* cur = ck[0];
  for (int s = 0; s < steps && rc == VR_OK; ++s) {
    NanoLeaf* dst;
    if (s + 1 == steps)
      dst = fin;
    else if ((s + 1) % c == 0)
      dst = ck[(s + 1) / c];
    else
      dst = seg[(s + 1) % c];
    step_forward(cur, dst);
    cur = dst;
  }
  if (rc == VR_OK) (void)hipEventRecord(ev[1], ctx->stream);
  // Terminal adjoint and loss; the per-voxel sums start at 0.
  if (rc == VR_OK) {
    hipLaunchKernelGGL(k_adj_leaf_active, dim3(grid_blocks(cap, 256)), dim3(256), 0, ctx->stream, phi0->d_leaves, L,
                       acnt);
    if ((e = vr_exclusive_scan(acnt, abase, (size_t)L, ascan, abase + cap, ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "scan");
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(3 * na, 256)), dim3(256), 0, ctx->stream, gacc, 3 * na);
    view.d_leaves = fin;
    hipLaunchKernelGGL(k_adj_terminal, dim3(leaf_blocks(L)), dim3(256), 0, ctx->stream, make_view(&view),
                       make_view(target), ob[0], leaf_loss);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, ctx->stream, leaf_loss, phi0->topo->d_count, loss_d);
    if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "terminal");
  }
  // Reverse, segment by segment; ob[bi] / far[bi] hold the last step's bricks.
  int bi = 0;
  for (int k = nseg - 1; k >= 0 && rc == VR_OK; --k) {
    const int s0 = k * c, s1 = (k + 1) * c < steps ? (k + 1) * c : steps;
    // Recompute states s0+1 .. s1-1 into seg[1..].
    NanoLeaf* prev = ck[k];
    for (int s = s0 + 1; s < s1 && rc == VR_OK; ++s) {
      step_forward(prev, seg[s - s0]);
      prev = seg[s - s0];
    }
    for (int s = s1 - 1; s >= s0 && rc == VR_OK; --s) {
      view.d_leaves = (s == s0) ? ck[k] : seg[s - s0];
      AdjStep P;
      P.phi = make_view(&view);
      P.vel = vv;
      P.ob_in = ob[bi];
      P.far_in = far[bi];
      P.ob_out = ob[bi ^ 1];
      P.far_out = far[bi ^ 1];
      P.grad = grad;
      P.abase = abase;
      P.gacc = gacc;
      P.gstride = (long long)na;
      P.dt = dt;
      P.tf0 = P.tf1 = P.tf2 = 1.0;
      P.separable = separable;
      if (scheme == VR_SCHEME_RK3)
        hipLaunchKernelGGL((k_adj<3, 0>), dim3(leaf_blocks(L)), dim3(ADJ_THREADS), 0, ctx->stream, P);
      else
        hipLaunchKernelGGL((k_adj<1, 0>), dim3(leaf_blocks(L)), dim3(ADJ_THREADS), 0, ctx->stream, P);
      // after the last backward step: the velocity gradient of all steps
      if (s == 0) {
        if (scheme == VR_SCHEME_RK3)
          hipLaunchKernelGGL((k_adj<3, 1>), dim3(leaf_blocks(L)), dim3(ADJ_THREADS), 0, ctx->stream, P);
        else
          hipLaunchKernelGGL((k_adj<1, 1>), dim3(leaf_blocks(L)), dim3(ADJ_THREADS), 0, ctx->stream, P);
      }
      if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "adjoint step");
      bi ^= 1;
    }
  }
  view.d_leaves = nullptr;
  if (rc == VR_OK) (void)hipEventRecord(ev[2], ctx->stream);
  vr_adjoint_stats st;
  std::memset(&st, 0, sizeof(st));
  if (rc == VR_OK && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) rc = vr_hip_fail(ctx, e, "adjoint sync");
  if (rc == VR_OK) {
    float f = 0.f, b = 0.f;
    (void)hipEventElapsedTime(&f, ev[0], ev[1]);
    (void)hipEventElapsedTime(&b, ev[1], ev[2]);
    st.forward_ms = f;
    st.adjoint_ms = b;
    st.steps = (uint64_t)steps;
    st.active_voxels = 0;
    if ((e = hipMemcpy(&st.loss, loss_d, sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "loss readback");
    if (rc == VR_OK && grad_out &&
        (e = hipMemcpy(grad_out, grad, sizeof(float) * 3 * nv3, hipMemcpyDeviceToHost)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "grad readback");
  }
  st.active_voxels = ac0;
  for (int k = 1; k < nseg; ++k)
    if (ck[k]) (void)hipFree(ck[k]);
  for (int j = 1; j < c; ++j)
    if (seg[j]) (void)hipFree(seg[j]);
  for (void* p : {(void*)fin, (void*)ob[0], (void*)ob[1], (void*)far[0], (void*)far[1], (void*)grad, (void*)leaf_loss,
                  (void*)loss_d, (void*)gacc, (void*)abase, (void*)acnt, (void*)ascan})
    if (p) (void)hipFree(p);
  for (int i = 0; i < 3; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (rc == VR_OK && stats) *stats = st;
  return rc;
}

// ------------------------------------------------------------------ several level sets
// vr_adjoint_rollout for `count` level sets at once: every forward step, the
// terminal adjoint and every backward step is one launch over the leaves of
// all grids.  States are kept as one leaf array over all grids (grid g at
// leaf_base[g]); the per-grid arguments of every launch (views of the state
// it reads, the buffers it writes) are laid out on the host in execution
// order and uploaded once.
extern "C" int vr_adjoint_rollout_batch(vr_ctx* ctx, int count, const vr_grid* const* phi0,
                                        const vr_velocity* const* vels, const double* dts, int scheme, int steps,
                                        int checkpoint_every, const vr_grid* const* targets, vr_adjoint_stats* stats,
                                        float* const* grads) {
  if (!ctx || count < 1 || !phi0 || !vels || !dts || !targets || steps < 1 || checkpoint_every < 1)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "bad adjoint arguments");
  if (scheme != VR_SCHEME_RK1 && scheme != VR_SCHEME_RK3)
    return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "scheme must be 1 or 3");
  const int G = count;
  std::vector<int> leaf_base((size_t)G + 1, 0);
  std::vector<size_t> goff((size_t)G + 1, 0);
  for (int g = 0; g < G; ++g) {
    if (!phi0[g] || !vels[g] || !targets[g]) return vr_fail(ctx, VR_ERR_INVALID_ARGUMENT, "null grid %d", g);
    if (vels[g]->kind != VK_DENSE || vels[g]->period > 0.0)
      return vr_fail(ctx, VR_ERR_UNSUPPORTED, "grid %d: the adjoint needs a static dense velocity field", g);
    if (phi0[g]->topo->has_empty)
      return vr_fail(ctx, VR_ERR_UNSUPPORTED, "grid %d: the adjoint needs leaves that all hold active voxels", g);
    leaf_base[g + 1] = leaf_base[g] + phi0[g]->topo->count;
    goff[g + 1] = goff[g] + 3 * (size_t)vels[g]->n * vels[g]->n * vels[g]->n;
    if ((long long)leaf_base[g + 1] > 0x7fffffffLL / 1000) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "batch too large");
  }
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return vr_hip_fail(ctx, e, "hipSetDevice");
  const int L = leaf_base[G];
  const size_t cap = (size_t)(L < 1 ? 1 : L);
  const int c = checkpoint_every;
  const int nseg = (steps + c - 1) / c;
  std::vector<NanoLeaf*> ck((size_t)nseg, nullptr), seg((size_t)c, nullptr);
  NanoLeaf* fin = nullptr;
  float *ob[2] = {nullptr, nullptr}, *far[2] = {nullptr, nullptr}, *grad = nullptr, *gacc = nullptr;
  double *leaf_loss = nullptr, *loss_d = nullptr;
  int *abase = nullptr, *acnt = nullptr, *ascan = nullptr;
  int* d_lb = nullptr;
  VelView* d_vel = nullptr;
  double* d_dts = nullptr;
  GridView *d_fin = nullptr, *d_tgt = nullptr, *d_fwd_in = nullptr;
  NanoLeaf** d_fwd_out = nullptr;
  AdjStep* d_bwd = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  int rc = VR_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc != VR_OK) return;
    if ((e = hipMalloc(p, bytes)) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipMalloc");
  };
  // active voxels of all grids: the stride of the per-voxel sums
  std::vector<uint64_t> act((size_t)G, 0);
  uint64_t atot = 0;
  for (int g = 0; g < G && rc == VR_OK; ++g) {
    uint64_t lc = 0;
    rc = vr_grid_counts(ctx, phi0[g], &lc, &act[g]);
    atot += act[g];
  }
  if (rc != VR_OK) return rc;
  if (atot > 0x7fffffffULL) return vr_fail(ctx, VR_ERR_UNSUPPORTED, "too many active voxels");
  const size_t na = atot ? (size_t)atot : 1;
  for (int k = 0; k < nseg; ++k) alloc((void**)&ck[k], sizeof(NanoLeaf) * cap);
  for (int j = 1; j < c; ++j) alloc((void**)&seg[j], sizeof(NanoLeaf) * cap);
  alloc((void**)&fin, sizeof(NanoLeaf) * cap);
  for (int i = 0; i < 2; ++i) {
    alloc((void**)&ob[i], sizeof(float) * 1000 * cap);
    alloc((void**)&far[i], sizeof(float) * 512 * cap);
  }
  alloc((void**)&grad, sizeof(float) * (goff[G] ? goff[G] : 1));
  alloc((void**)&gacc, sizeof(float) * 3 * na);
  alloc((void**)&abase, sizeof(int) * (cap + 1));
  alloc((void**)&acnt, sizeof(int) * cap);
  alloc((void**)&ascan, sizeof(int) * (vr_scan_scratch_ints(cap) + 1));
  alloc((void**)&leaf_loss, sizeof(double) * cap);
  alloc((void**)&loss_d, sizeof(double) * G);
  alloc((void**)&d_lb, sizeof(int) * (G + 1));
  alloc((void**)&d_vel, sizeof(VelView) * G);
  alloc((void**)&d_dts, sizeof(double) * G);
  alloc((void**)&d_fin, sizeof(GridView) * G);
  alloc((void**)&d_tgt, sizeof(GridView) * G);
  for (int i = 0; i < 3 && rc == VR_OK; ++i)
    if ((e = hipEventCreate(&ev[i])) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipEventCreate");
  // Views of grid g over a state buffer (phi0's topology).
  std::vector<std::unique_ptr<vr_grid>> views;
  std::vector<int> separable((size_t)G, 1);
  std::vector<VelView> vv((size_t)G);
  for (int g = 0; g < G; ++g) {
    std::unique_ptr<vr_grid> v(new vr_grid());
    copy_header_into(v.get(), phi0[g]);
    v->topo = phi0[g]->topo;
    views.push_back(std::move(v));
    vv[g] = make_vel_view(vels[g]);
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q)
        if (r == q ? !(phi0[g]->mat[r][q] > 0.0) : phi0[g]->mat[r][q] != 0.0) separable[g] = 0;
  }
  auto view_of = [&](int g, NanoLeaf* state) {
    views[g]->d_leaves = state + leaf_base[g];
    const GridView gv = make_view(views[g].get());
    views[g]->d_leaves = nullptr;  // not owned
    return gv;
  };
  // The launch sequence, laid out on the host in execution order.
  std::vector<GridView> fwd_in, fin_v, tgt_v;
  std::vector<NanoLeaf*> fwd_out;
  std::vector<AdjStep> bwd;
  struct Op {
    int kind;  // 0 forward, 1 backward
    int index;
  };
  std::vector<Op> ops;
  auto add_forward = [&](NanoLeaf* src, NanoLeaf* dst) {
    ops.push_back({0, (int)(fwd_in.size() / G)});
    for (int g = 0; g < G; ++g) {
      fwd_in.push_back(view_of(g, src));
      fwd_out.push_back(dst + leaf_base[g]);
    }
  };
  NanoLeaf* cur = ck[0];
  for (int s = 0; s < steps; ++s) {
    NanoLeaf* dst = s + 1 == steps ? fin : ((s + 1) % c == 0 ? ck[(s + 1) / c] : seg[(s + 1) % c]);
    add_forward(cur, dst);
    cur = dst;
  }
  const size_t n_forward_ops = ops.size();
  for (int g = 0; g < G; ++g) {
    fin_v.push_back(view_of(g, fin));
    tgt_v.push_back(make_view(targets[g]));
  }
  int bi = 0;
  for (int k = nseg - 1; k >= 0; --k) {
    const int s0 = k * c, s1 = (k + 1) * c < steps ? (k + 1) * c : steps;
    NanoLeaf* prev = ck[k];
    for (int s = s0 + 1; s < s1; ++s) {
      add_forward(prev, seg[s - s0]);
      prev = seg[s - s0];
    }
    for (int s = s1 - 1; s >= s0; --s) {
      NanoLeaf* state = (s == s0) ? ck[k] : seg[s - s0];
      ops.push_back({1, (int)(bwd.size() / G)});
      for (int g = 0; g < G; ++g) {
        const size_t lb = (size_t)leaf_base[g];
        AdjStep P;
        P.phi = view_of(g, state);
        P.vel = vv[g];
        P.ob_in = ob[bi] + lb * 1000;
        P.far_in = far[bi] + lb * 512;
        P.ob_out = ob[bi ^ 1] + lb * 1000;
        P.far_out = far[bi ^ 1] + lb * 512;
        P.grad = grad + goff[g];
        P.abase = abase + lb;
        P.gacc = gacc;
        P.gstride = (long long)na;
        P.dt = dts[g];
        P.tf0 = P.tf1 = P.tf2 = 1.0;
        P.separable = separable[g];
        bwd.push_back(P);
      }
      bi ^= 1;
    }
  }
  alloc((void**)&d_fwd_in, sizeof(GridView) * (fwd_in.size() ? fwd_in.size() : 1));
  alloc((void**)&d_fwd_out, sizeof(NanoLeaf*) * (fwd_out.size() ? fwd_out.size() : 1));
  alloc((void**)&d_bwd, sizeof(AdjStep) * (bwd.size() ? bwd.size() : 1));
  auto h2d = [&](void* dst, const void* src, size_t bytes) {
    if (rc == VR_OK && bytes && (e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "hipMemcpyAsync");
  };
  std::vector<double> dtv(dts, dts + G);
  h2d(d_lb, leaf_base.data(), sizeof(int) * (G + 1));
  h2d(d_vel, vv.data(), sizeof(VelView) * G);
  h2d(d_dts, dtv.data(), sizeof(double) * G);
  h2d(d_fin, fin_v.data(), sizeof(GridView) * G);
  h2d(d_tgt, tgt_v.data(), sizeof(GridView) * G);
  h2d(d_fwd_in, fwd_in.data(), sizeof(GridView) * fwd_in.size());
  h2d(d_fwd_out, fwd_out.data(), sizeof(NanoLeaf*) * fwd_out.size());
  h2d(d_bwd, bwd.data(), sizeof(AdjStep) * bwd.size());
  for (int g = 0; g < G && rc == VR_OK; ++g)
    if (phi0[g]->topo->count &&
        (e = hipMemcpyAsync(ck[0] + leaf_base[g], phi0[g]->d_leaves, sizeof(NanoLeaf) * phi0[g]->topo->count,
                            hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "copy");
  if (rc == VR_OK) {
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(goff[G], 256)), dim3(256), 0, ctx->stream, grad, goff[G]);
    for (int i = 0; i < 2; ++i)
      hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(512 * cap, 256)), dim3(256), 0, ctx->stream, far[i], 512 * cap);
  }
  const int nw = vr_advect_waves(L, ctx->cus) == 1 ? 1 : 4;
  auto launch_forward = [&](int index) {
    const GridView* in = d_fwd_in + (size_t)index * G;
    NanoLeaf* const* out = d_fwd_out + (size_t)index * G;
    const int blocks = leaf_blocks((size_t)L);
    if (scheme == VR_SCHEME_RK3) {
      if (nw == 1)
        hipLaunchKernelGGL((k_adj_advect_multi<3, 1>), dim3(blocks), dim3(64), 0, ctx->stream, in, out, d_vel, d_dts,
                           d_lb, G, L);
      else
        hipLaunchKernelGGL((k_adj_advect_multi<3, 4>), dim3(blocks), dim3(256), 0, ctx->stream, in, out, d_vel, d_dts,
                           d_lb, G, L);
    } else {
      if (nw == 1)
        hipLaunchKernelGGL((k_adj_advect_multi<1, 1>), dim3(blocks), dim3(64), 0, ctx->stream, in, out, d_vel, d_dts,
                           d_lb, G, L);
      else
        hipLaunchKernelGGL((k_adj_advect_multi<1, 4>), dim3(blocks), dim3(256), 0, ctx->stream, in, out, d_vel, d_dts,
                           d_lb, G, L);
    }
  };
  if (rc == VR_OK) (void)hipEventRecord(ev[0], ctx->stream);
  for (size_t o = 0; o < n_forward_ops && rc == VR_OK; ++o) launch_forward(ops[o].index);
  if (rc == VR_OK) {
    (void)hipEventRecord(ev[1], ctx->stream);
    // per-voxel sums: offsets over the leaves of all grids, zeroed
    hipLaunchKernelGGL(k_adj_leaf_active, dim3(grid_blocks(cap, 256)), dim3(256), 0, ctx->stream, ck[0], L, acnt);
    if ((e = vr_exclusive_scan(acnt, abase, (size_t)L, ascan, abase + cap, ctx->stream)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "scan");
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(3 * na, 256)), dim3(256), 0, ctx->stream, gacc, 3 * na);
    hipLaunchKernelGGL(k_adj_terminal_multi, dim3(leaf_blocks((size_t)L)), dim3(256), 0, ctx->stream, d_fin, d_tgt,
                       ob[0], leaf_loss, d_lb, G, L);
    for (int g = 0; g < G; ++g)
      hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, ctx->stream, leaf_loss + leaf_base[g], phi0[g]->topo->d_count,
                         loss_d + g);
  }
  for (size_t o = n_forward_ops; o < ops.size() && rc == VR_OK; ++o) {
    if (ops[o].kind == 0) {
      launch_forward(ops[o].index);
    } else {
      const AdjStep* P = d_bwd + (size_t)ops[o].index * G;
      if (scheme == VR_SCHEME_RK3)
        hipLaunchKernelGGL((k_adj_multi<3, 0>), dim3(leaf_blocks((size_t)L)), dim3(ADJ_THREADS), 0, ctx->stream, P,
                           d_lb, G, L);
      else
        hipLaunchKernelGGL((k_adj_multi<1, 0>), dim3(leaf_blocks((size_t)L)), dim3(ADJ_THREADS), 0, ctx->stream, P,
                           d_lb, G, L);
      // after the last backward step (state 0): the velocity gradient of all steps
      if (o + 1 == ops.size()) {
        if (scheme == VR_SCHEME_RK3)
          hipLaunchKernelGGL((k_adj_multi<3, 1>), dim3(leaf_blocks((size_t)L)), dim3(ADJ_THREADS), 0, ctx->stream, P,
                             d_lb, G, L);
        else
          hipLaunchKernelGGL((k_adj_multi<1, 1>), dim3(leaf_blocks((size_t)L)), dim3(ADJ_THREADS), 0, ctx->stream, P,
                             d_lb, G, L);
      }
    }
    if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "adjoint launch");
  }
  if (rc == VR_OK) (void)hipEventRecord(ev[2], ctx->stream);
  if (rc == VR_OK && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) rc = vr_hip_fail(ctx, e, "adjoint sync");
  std::vector<double> loss((size_t)G, 0.0);
  float fms = 0.f, bms = 0.f;
  if (rc == VR_OK) {
    (void)hipEventElapsedTime(&fms, ev[0], ev[1]);
    (void)hipEventElapsedTime(&bms, ev[1], ev[2]);
    if ((e = hipMemcpy(loss.data(), loss_d, sizeof(double) * G, hipMemcpyDeviceToHost)) != hipSuccess)
      rc = vr_hip_fail(ctx, e, "loss readback");
  }
  for (int g = 0; g < G && rc == VR_OK; ++g) {
    if (grads && grads[g] &&
        (e = hipMemcpy(grads[g], grad + goff[g], sizeof(float) * (goff[g + 1] - goff[g]), hipMemcpyDeviceToHost)) !=
            hipSuccess)
      rc = vr_hip_fail(ctx, e, "grad readback");
    if (rc == VR_OK && stats) {
      const uint64_t ac = act[g];
      vr_adjoint_stats& st = stats[g];
      std::memset(&st, 0, sizeof(st));
      st.loss = loss[g];
      st.forward_ms = fms;
      st.adjoint_ms = bms;
      st.steps = (uint64_t)steps;
      st.active_voxels = ac;
    }
  }
  for (NanoLeaf* p : ck)
    if (p) (void)hipFree(p);
  for (int j = 1; j < c; ++j)
    if (seg[j]) (void)hipFree(seg[j]);
  for (void* p : {(void*)fin, (void*)ob[0], (void*)ob[1], (void*)far[0], (void*)far[1], (void*)grad, (void*)leaf_loss,
                  (void*)loss_d, (void*)d_lb, (void*)d_vel, (void*)d_dts, (void*)d_fin, (void*)d_tgt, (void*)d_fwd_in,
                  (void*)d_fwd_out, (void*)d_bwd, (void*)gacc, (void*)abase, (void*)acnt, (void*)ascan})
    if (p) (void)hipFree(p);
  for (int i = 0; i < 3; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  return rc;
}
