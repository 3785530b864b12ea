// Differentiable rollout: forward advection steps through a static dense
// velocity field with checkpoints, then the adjoint (reverse) pass.
//
// No reference counterpart (the reference has no differentiation).  The
// step is phi_{n+1}(x_i) = S(phi_n, p_i(U)), S the trilinear sample
// (src/tools/interpolation.rs:11-48) and p_i the RK1 or RK3 departure point
// through the dense field U.  Only active voxels are variables; inactive and
// background values are constants.  Per backward step and per leaf (one
// block of 4 waves):
//   - the adjoint lambda_{n+1} of each active voxel is scattered with the
//     trilinear weights into per-wave LDS bricks (leaf + one-voxel halo),
//     summed in wave order into an outbox; a merge kernel sums, for each
//     voxel, the entries of the (up to 8) leaves whose brick covers it in a
//     fixed order, so lambda_n is deterministic;
//   - lambda * dS/dp is pulled back through the departure point (and, for
//     RK3, through the stage positions with the field's Jacobian) onto the
//     velocity nodes, pre-reduced in an LDS window and added to dL/dU with
//     global atomics.
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
  const float* lam_next;
  float* outbox;   // [cap][1000]
  float* lam_far;  // [cap][512], corners beyond the brick (large steps)
  float* grad;     // 3 n^3
  double dt, tf0, tf1, tf2;
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
// else global.  Lanes visit the 8 nodes in a lane-rotated order: voxels of
// one wave mostly share a velocity cell, and the rotation makes the lanes of
// one atomic instruction hit different nodes instead of serialising on one.
__device__ __forceinline__ void vel_scatter(const AdjStep& P, float* vwin, const int wb[3], const double p[3],
                                            const double coef[3]) {
  int b[3];
  double f[3];
  dense_weights(P.vel, p, b, f);
  const int n = P.vel.n;
  const size_t n3 = (size_t)n * n * n;
  const int lane = threadIdx.x & 63;
  const int rot = lane + (lane >> 3);
  for (int s = 0; s < 8; ++s) {
    const int q = (s + rot) & 7;
    const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
    const int i = b[0] + dx, j = b[1] + dy, k = b[2] + dz;
    if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n || (unsigned)k >= (unsigned)n) continue;
    const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
    const int wi = i - wb[0], wj = j - wb[1], wk = k - wb[2];
    const bool in = (unsigned)wi < VW && (unsigned)wj < VW && (unsigned)wk < VW;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float val = (float)(coef[c] * w);
      if (in)
        lds_add(&vwin[c * VW3 + (wi * VW + wj) * VW + wk], val);
      else
        atomicAdd(&P.grad[c * n3 + ((size_t)i * n + j) * n + k], val);
    }
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

#define ADJ_WAVES 4
#define ADJ_THREADS (64 * ADJ_WAVES)

// One backward step, one leaf per block of ADJ_WAVES waves.  Each wave
// scatters its voxels' lambda into its own LDS brick (a wave's LDS atomics
// apply in a fixed order), and the outbox entry is the sum of the wave
// bricks in wave order, so lambda_n is the same on every run.
template <int SCHEME>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_step(AdjStep P) {
  __shared__ float brick[1000];
  __shared__ float acc[ADJ_WAVES][1000];
  __shared__ float vwin[3 * VW3];
  __shared__ uint16_t list[512];
  __shared__ int nb[VR_NBR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const GridView& g = P.phi;
  const int n = *g.count;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const double tf[3] = {P.tf0, P.tf1, P.tf2};
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = g.leaves[li];
    if (tid < VR_NBR) nb[tid] = g.nbr[(size_t)li * VR_NBR + tid];
    int total = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const uint64_t m = L.mask[x];
      if ((x % ADJ_WAVES) == wave && ((m >> lane) & 1ull)) list[total + __popcll(m & below)] = (uint16_t)(x * 64 + lane);
      total += __popcll(m);
    }
    for (int i = tid; i < ADJ_WAVES * 1000; i += ADJ_THREADS) (&acc[0][0])[i] = 0.0f;
    for (int i = tid; i < 3 * VW3; i += ADJ_THREADS) vwin[i] = 0.0f;
    const int ox = L.origin[0], oy = L.origin[1], oz = L.origin[2];
    int wb[3];
    {
      double ow[3];
      index_to_world(g.mat, ox, oy, oz, ow);
      for (int a = 0; a < 3; ++a) wb[a] = f64_to_i32(floor(ow[a] * P.vel.inv_h)) - 2;
    }
    stage_interior<ADJ_THREADS>(g, li, brick, tid);
    __syncthreads();
    stage_halo<ADJ_THREADS>(g, nb, brick, tid);
    __syncthreads();
    float* const wacc = acc[wave];
    for (int ii = tid; ii < total; ii += ADJ_THREADS) {
      const int off = list[ii];
      const float lam = P.lam_next[(size_t)li * 512 + off];
      if (lam == 0.0f) continue;
      const int x = off >> 6, y = (off >> 3) & 7, z = off & 7;
      double xw[3], y1[3], y2[3], p[3], u[3];
      index_to_world(g.mat, ox + x, oy + y, oz + z, xw);
      vel_sample<VK_DENSE>(P.vel, xw, u);
      for (int a = 0; a < 3; ++a) y1[a] = xw[a] - P.dt * (u[a] * tf[0]);
      if (SCHEME == 1) {
        for (int a = 0; a < 3; ++a) p[a] = y1[a];
      } else {
        vel_sample<VK_DENSE>(P.vel, y1, u);
        for (int a = 0; a < 3; ++a) y2[a] = 0.75 * xw[a] + 0.25 * (y1[a] - P.dt * (u[a] * tf[1]));
        vel_sample<VK_DENSE>(P.vel, y2, u);
        for (int a = 0; a < 3; ++a) p[a] = (1.0 / 3.0) * xw[a] + (2.0 / 3.0) * (y2[a] - P.dt * (u[a] * tf[2]));
      }
      const double ix = p[0] * g.inv_vs, iy = p[1] * g.inv_vs, iz = p[2] * g.inv_vs;
      const double flx = floor(ix), fly = floor(iy), flz = floor(iz);
      const double fx = ix - flx, fy = iy - fly, fz = iz - flz;
      const int i0 = f64_to_i32(flx), j0 = f64_to_i32(fly), k0 = f64_to_i32(flz);
      const int bi = i0 - ox, bj = j0 - oy, bk = k0 - oz;
      const bool inb = bi >= -1 && bi <= 7 && bj >= -1 && bj <= 7 && bk >= -1 && bk <= 7;
      double c[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
        const double w = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
        if (inb) {
          const int bidx = (bi + 1 + dx) * 100 + (bj + 1 + dy) * 10 + (bk + 1 + dz);
          c[q] = (double)brick[bidx];
          lds_add(&wacc[bidx], (float)((double)lam * w));
        } else {
          const int cx = i0 + dx, cy = j0 + dy, cz = k0 + dz;
          const int tl = find_leaf(g, cx, cy, cz);
          c[q] = tl < 0 ? (double)g.background : (double)g.leaves[tl].values[voxel_offset(cx, cy, cz)];
          if (tl >= 0) atomicAdd(&P.lam_far[(size_t)tl * 512 + voxel_offset(cx, cy, cz)], (float)((double)lam * w));
        }
      }
      const double dfx = (c[1] - c[0]) * (1 - fy) * (1 - fz) + (c[3] - c[2]) * fy * (1 - fz) +
                         (c[5] - c[4]) * (1 - fy) * fz + (c[7] - c[6]) * fy * fz;
      const double dfy = (c[2] - c[0]) * (1 - fx) * (1 - fz) + (c[3] - c[1]) * fx * (1 - fz) +
                         (c[6] - c[4]) * (1 - fx) * fz + (c[7] - c[5]) * fx * fz;
      const double dfz = (c[4] - c[0]) * (1 - fx) * (1 - fy) + (c[5] - c[1]) * fx * (1 - fy) +
                         (c[6] - c[2]) * (1 - fx) * fy + (c[7] - c[3]) * fx * fy;
      const double gp[3] = {(double)lam * dfx * g.inv_vs, (double)lam * dfy * g.inv_vs, (double)lam * dfz * g.inv_vs};
      if (SCHEME == 1) {
        const double coef[3] = {-P.dt * tf[0] * gp[0], -P.dt * tf[0] * gp[1], -P.dt * tf[0] * gp[2]};
        vel_scatter(P, vwin, wb, xw, coef);
      } else {
        double jt[3], gy2[3], gy1[3];
        dense_jt(P.vel, y2, gp, jt);
        for (int a = 0; a < 3; ++a) gy2[a] = (2.0 / 3.0) * gp[a] - (2.0 / 3.0) * P.dt * tf[2] * jt[a];
        const double c3[3] = {-(2.0 / 3.0) * P.dt * tf[2] * gp[0], -(2.0 / 3.0) * P.dt * tf[2] * gp[1],
                              -(2.0 / 3.0) * P.dt * tf[2] * gp[2]};
        vel_scatter(P, vwin, wb, y2, c3);
        dense_jt(P.vel, y1, gy2, jt);
        for (int a = 0; a < 3; ++a) gy1[a] = 0.25 * gy2[a] - 0.25 * P.dt * tf[1] * jt[a];
        const double c2[3] = {-0.25 * P.dt * tf[1] * gy2[0], -0.25 * P.dt * tf[1] * gy2[1],
                              -0.25 * P.dt * tf[1] * gy2[2]};
        vel_scatter(P, vwin, wb, y1, c2);
        const double c1[3] = {-P.dt * tf[0] * gy1[0], -P.dt * tf[0] * gy1[1], -P.dt * tf[0] * gy1[2]};
        vel_scatter(P, vwin, wb, xw, c1);
      }
    }
    __syncthreads();
    float* ob = P.outbox + (size_t)li * 1000;
    for (int i = tid; i < 1000; i += ADJ_THREADS) {
      float v = acc[0][i];
#pragma unroll
      for (int w = 1; w < ADJ_WAVES; ++w) v += acc[w][i];
      ob[i] = v;
    }
    const int nv = P.vel.n;
    const size_t n3 = (size_t)nv * nv * nv;
    for (int i = tid; i < 3 * VW3; i += ADJ_THREADS) {
      const float v = vwin[i];
      if (v == 0.0f) continue;
      const int c = i / VW3, r = i % VW3;
      const int gi = wb[0] + r / (VW * VW), gj = wb[1] + (r / VW) % VW, gk = wb[2] + r % VW;
      atomicAdd(&P.grad[c * n3 + ((size_t)gi * nv + gj) * nv + gk], v);
    }
    __syncthreads();
  }
}

// lambda_n(y) for each active voxel y: the outbox entries of the leaves whose
// brick covers y, summed in neighbour order, plus the far-corner buffer.  A
// neighbour at offset (dx, dy, dz) covers y only where y's local coordinate
// is 0 (for d = -1) or 7 (for d = +1) on every axis with d != 0, so at most 8
// bricks are visited, in ascending neighbour order.
__global__ __launch_bounds__(256) void k_adj_merge(GridView g, const float* outbox, float* lam_far, float* lam) {
  __shared__ int nb[VR_NBR];
  const int tid = threadIdx.x;
  const int n = *g.count;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    if (tid < VR_NBR) nb[tid] = g.nbr[(size_t)li * VR_NBR + tid];
    __syncthreads();
    const NanoLeaf& L = g.leaves[li];
    for (int off = tid; off < 512; off += 256) {
      const int x = off >> 6, y = (off >> 3) & 7, z = off & 7;
      float s = 0.0f;
      if ((L.mask[x] >> (off & 63)) & 1ull) {
        for (int dx = (x == 0 ? -1 : 0); dx <= (x == 7 ? 1 : 0); ++dx)
          for (int dy = (y == 0 ? -1 : 0); dy <= (y == 7 ? 1 : 0); ++dy)
            for (int dz = (z == 0 ? -1 : 0); dz <= (z == 7 ? 1 : 0); ++dz) {
              const int src = nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
              if (src < 0) continue;
              const int lx = x - 8 * dx, ly = y - 8 * dy, lz = z - 8 * dz;
              s += outbox[(size_t)src * 1000 + (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1)];
            }
        s += lam_far[(size_t)li * 512 + off];
      }
      lam[(size_t)li * 512 + off] = s;
      lam_far[(size_t)li * 512 + off] = 0.0f;
    }
    __syncthreads();
  }
}

// lambda_K = phi_K - target at active voxels (0 elsewhere); per-leaf loss in f64.
__global__ __launch_bounds__(64) void k_adj_terminal(GridView g, GridView target, float* lam, double* leaf_loss) {
  const int lane = threadIdx.x;
  const int n = *g.count;
  const int y = lane >> 3, z = lane & 7;
  for (int li = blockIdx.x; li < n; li += gridDim.x) {
    const NanoLeaf& L = g.leaves[li];
    double part = 0.0;
    for (int x = 0; x < 8; ++x) {
      const int off = x * 64 + lane;
      double d = 0.0;
      if ((L.mask[x] >> lane) & 1ull) {
        const float t = grid_get(target, L.origin[0] + x, L.origin[1] + y, L.origin[2] + z);
        d = (double)L.values[off] - (double)t;
        part += 0.5 * d * d;
      }
      lam[(size_t)li * 512 + off] = (float)d;
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
    if (lane == 0) leaf_loss[li] = part;
  }
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
  float *lam_a = nullptr, *lam_b = nullptr, *outbox = nullptr, *far = nullptr, *grad = nullptr;
  double *leaf_loss = nullptr, *loss_d = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  int rc = VR_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc != VR_OK) return;
    if ((e = hipMalloc(p, bytes)) != hipSuccess) rc = vr_hip_fail(ctx, e, "hipMalloc");
  };
  ck[0] = phi0->d_leaves;
  for (int k = 1; k < nseg; ++k) alloc((void**)&ck[k], sizeof(NanoLeaf) * cap);
  for (int j = 1; j < c; ++j) alloc((void**)&seg[j], sizeof(NanoLeaf) * cap);
  alloc((void**)&fin, sizeof(NanoLeaf) * cap);
  alloc((void**)&lam_a, sizeof(float) * 512 * cap);
  alloc((void**)&lam_b, sizeof(float) * 512 * cap);
  alloc((void**)&outbox, sizeof(float) * 1000 * cap);
  alloc((void**)&far, sizeof(float) * 512 * cap);
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
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_blocks(512 * (size_t)cap, 256)), dim3(256), 0, ctx->stream, far,
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
  // Terminal adjoint and loss.
  if (rc == VR_OK) {
    view.d_leaves = fin;
    hipLaunchKernelGGL(k_adj_terminal, dim3(grid_blocks(L, 1)), dim3(64), 0, ctx->stream, make_view(&view),
                       make_view(target), lam_a, leaf_loss);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, ctx->stream, leaf_loss, phi0->topo->d_count, loss_d);
    if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "terminal");
  }
  // Reverse, segment by segment.
  float* lam_next = lam_a;
  float* lam_cur = lam_b;
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
      P.lam_next = lam_next;
      P.outbox = outbox;
      P.lam_far = far;
      P.grad = grad;
      P.dt = dt;
      P.tf0 = P.tf1 = P.tf2 = 1.0;
      if (scheme == VR_SCHEME_RK3)
        hipLaunchKernelGGL(k_adj_step<3>, dim3(grid_blocks(L, 1)), dim3(ADJ_THREADS), 0, ctx->stream, P);
      else
        hipLaunchKernelGGL(k_adj_step<1>, dim3(grid_blocks(L, 1)), dim3(ADJ_THREADS), 0, ctx->stream, P);
      hipLaunchKernelGGL(k_adj_merge, dim3(grid_blocks(L, 1)), dim3(256), 0, ctx->stream, P.phi, outbox, far, lam_cur);
      if ((e = hipGetLastError()) != hipSuccess) rc = vr_hip_fail(ctx, e, "adjoint step");
      std::swap(lam_next, lam_cur);
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
  if (rc == VR_OK) {
    uint64_t lc = 0, ac = 0;
    rc = vr_grid_counts(ctx, phi0, &lc, &ac);
    st.active_voxels = ac;
  }
  for (int k = 1; k < nseg; ++k)
    if (ck[k]) (void)hipFree(ck[k]);
  for (int j = 1; j < c; ++j)
    if (seg[j]) (void)hipFree(seg[j]);
  for (void* p : {(void*)fin, (void*)lam_a, (void*)lam_b, (void*)outbox, (void*)far, (void*)grad, (void*)leaf_loss,
                  (void*)loss_d})
    if (p) (void)hipFree(p);
  for (int i = 0; i < 3; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (rc == VR_OK && stats) *stats = st;
  return rc;
}
