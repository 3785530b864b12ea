// volrus oracle, builder-defined parts — TEST INFRASTRUCTURE ONLY.
//
// The reference has no RK3 back-trace, no analytic or dense velocity, no
// adjoint and no slot/noise/union fixture generators (its level-set advection
// is RK1 only, src/tools/advection.rs:32-56).  This file states those parts
// once, in plain double/float arithmetic, so that the device path can be
// checked against them.  Each routine is pinned by tests on its own
// (closed forms, reductions to the reference's RK1, finite differences).
#pragma once

#include "volrus_oracle.hpp"

namespace vo {

// ---------------------------------------------------------------- SplitMix64
// Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014 (the finaliser constants are the published ones).
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

// ---------------------------------------------------------------- sin/cos of pi*x
// Deterministic sinpi/cospi: exact reduction to f in [-1/4, 1/4], then the
// Taylor series of sin(pi f) and cos(pi f) to f^15 / f^14 (truncation below
// 4e-16 and 1e-15), Horner with explicit fused multiply-adds (one IEEE
// rounding each, std::fma), so any IEEE machine gives the same bits.
inline void sincospi_det(double x, double* s_out, double* c_out) {
  const double n = std::nearbyint(2.0 * x);
  const double f = x - 0.5 * n;
  const double f2 = f * f;
  double s = -2.1915353447830217e-05;
  s = std::fma(s, f2, 0.00046630280576761255);
  s = std::fma(s, f2, -0.0073704309457143504);
  s = std::fma(s, f2, 0.08214588661112823);
  s = std::fma(s, f2, -0.5992645293207921);
  s = std::fma(s, f2, 2.5501640398773455);
  s = std::fma(s, f2, -5.16771278004997);
  s = std::fma(s, f2, 3.141592653589793);
  s = s * f;
  double c = -0.0001046381049248457;
  c = std::fma(c, f2, 0.0019295743094039231);
  c = std::fma(c, f2, -0.02580689139001406);
  c = std::fma(c, f2, 0.2353306303588932);
  c = std::fma(c, f2, -1.3352627688545895);
  c = std::fma(c, f2, 4.0587121264167685);
  c = std::fma(c, f2, -4.934802200544679);
  c = std::fma(c, f2, 1.0);
  const long q = ((long)n) & 3;
  switch (q) {
    case 0: *s_out = s; *c_out = c; break;
    case 1: *s_out = c; *c_out = -s; break;
    case 2: *s_out = -s; *c_out = -c; break;
    default: *s_out = -c; *c_out = s; break;
  }
}

// ---------------------------------------------------------------- analytic velocity
// Builder-defined fields, evaluated in double at world positions.
//   UNIFORM: u = (a0, a1, a2)
//   RIGID:   u = w x (x - c), w = (a0, a1, a2), c = (a3, a4, a5)
//   VORTEX:  u = (sin^2(pi x) sin(2 pi y), -sin^2(pi y) sin(2 pi x), 0)
// A time factor cospi(t / period) multiplies every field when period > 0.
// Numbered like the product's VR_VEL_* kinds.
enum AnalyticKind { AK_UNIFORM = 2, AK_RIGID = 3, AK_VORTEX = 4 };

struct AnalyticVelocity {
  int kind = AK_UNIFORM;
  double a[6] = {0, 0, 0, 0, 0, 0};
  void sample(const double p[3], double out[3]) const {
    if (kind == AK_UNIFORM) {
      out[0] = a[0]; out[1] = a[1]; out[2] = a[2];
    } else if (kind == AK_RIGID) {
      const double rx = p[0] - a[3], ry = p[1] - a[4], rz = p[2] - a[5];
      out[0] = a[1] * rz - a[2] * ry;
      out[1] = a[2] * rx - a[0] * rz;
      out[2] = a[0] * ry - a[1] * rx;
    } else {
      double sx, cx, sy, cy;
      sincospi_det(p[0], &sx, &cx);
      sincospi_det(p[1], &sy, &cy);
      out[0] = (sx * sx) * (2.0 * sy * cy);
      out[1] = -((sy * sy) * (2.0 * sx * cx));
      out[2] = 0.0;
    }
  }
};

inline double time_factor(double period, double t) {
  if (!(period > 0.0)) return 1.0;
  double s, c;
  sincospi_det(t / period, &s, &c);
  return c;
}

// ---------------------------------------------------------------- dense velocity
// Builder-defined: three n^3 float arrays over nodes (i, j, k) at world
// (i h, j h, k h), index (i * n + j) * n + k, value 0 outside [0, n)^3.
// Sampling is exactly src/tools/interpolation.rs:11-48 applied to a grid of
// voxel size h with background 0.
struct DenseVelocity {
  int n = 0;
  double h = 1.0;
  std::vector<float> u[3];
  float at(int c, int32_t i, int32_t j, int32_t k) const {
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return 0.0f;
    return u[c][((size_t)i * n + j) * n + k];
  }
  float sample_comp(int c, const double w[3]) const {
    const double inv = 1.0 / h;
    const double idx[3] = {w[0] * inv, w[1] * inv, w[2] * inv};
    const double ix = std::floor(idx[0]), iy = std::floor(idx[1]), iz = std::floor(idx[2]);
    const float fx = (float)(idx[0] - ix), fy = (float)(idx[1] - iy), fz = (float)(idx[2] - iz);
    const int32_t i = f64_to_i32(ix), j = f64_to_i32(iy), k = f64_to_i32(iz);
    const float c000 = at(c, i, j, k), c100 = at(c, i + 1, j, k), c010 = at(c, i, j + 1, k),
                c110 = at(c, i + 1, j + 1, k), c001 = at(c, i, j, k + 1), c101 = at(c, i + 1, j, k + 1),
                c011 = at(c, i, j + 1, k + 1), c111 = at(c, i + 1, j + 1, k + 1);
    const float c00 = c000 * (1.0f - fx) + c100 * fx;
    const float c01 = c001 * (1.0f - fx) + c101 * fx;
    const float c10 = c010 * (1.0f - fx) + c110 * fx;
    const float c11 = c011 * (1.0f - fx) + c111 * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy;
    const float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
  }
  void sample(const double p[3], double out[3]) const {
    for (int c = 0; c < 3; ++c) out[c] = (double)sample_comp(c, p);
  }
  // Trilinear weights of node (i + a, j + b, k + d) at p, in double.
  void weights(const double p[3], int32_t base[3], double w[8]) const {
    const double inv = 1.0 / h;
    double f[3];
    for (int a = 0; a < 3; ++a) {
      const double idx = p[a] * inv, fl = std::floor(idx);
      base[a] = f64_to_i32(fl);
      f[a] = idx - fl;
    }
    for (int q = 0; q < 8; ++q) {
      const int a = q & 1, b = (q >> 1) & 1, d = (q >> 2) & 1;
      w[q] = (a ? f[0] : 1.0 - f[0]) * (b ? f[1] : 1.0 - f[1]) * (d ? f[2] : 1.0 - f[2]);
    }
  }
};

// ---------------------------------------------------------------- RK1 / RK3 step
// Builder-defined generalisation of src/tools/advection.rs:32-56.  Velocity
// at stage s is sample(p) * tf[s]; RK1 uses tf[0] (time t + dt).  RK3 is the
// SSP (Shu-Osher) scheme applied to the backward characteristic
// dX/ds = -u(X, t + dt - s):
//   y1 = x - dt u(x, t+dt)
//   y2 = 3/4 x + 1/4 (y1 - dt u(y1, t))
//   y3 = 1/3 x + 2/3 (y2 - dt u(y2, t + dt/2))
// With tf = {1, 1, 1} and RK1 this is the reference step bit for bit.
enum Scheme { SCHEME_RK1 = 1, SCHEME_RK3 = 3 };

template <class T, class Vel>
void departure_point(const Grid<T>& grid, const Vel& vel, Coord c, double dt, int scheme, const double tf[3],
                     double out[3]) {
  double x[3], v[3];
  grid.affine.index_to_world(c, x);
  vel.sample(x, v);
  if (scheme == SCHEME_RK1) {
    for (int a = 0; a < 3; ++a) out[a] = x[a] - dt * (v[a] * tf[0]);
    return;
  }
  double y1[3], y2[3];
  for (int a = 0; a < 3; ++a) y1[a] = x[a] - dt * (v[a] * tf[0]);
  vel.sample(y1, v);
  for (int a = 0; a < 3; ++a) y2[a] = 0.75 * x[a] + 0.25 * (y1[a] - dt * (v[a] * tf[1]));
  vel.sample(y2, v);
  for (int a = 0; a < 3; ++a) out[a] = (1.0 / 3.0) * x[a] + (2.0 / 3.0) * (y2[a] - dt * (v[a] * tf[2]));
}

template <class T, class Vel>
Grid<T> advect_scheme(const Grid<T>& grid, const Vel& vel, double dt, int scheme, const double tf[3]) {
  Grid<T> out = Grid<T>::with_affine(grid.background(), grid.affine);
  out.name = grid.name;
  out.grid_class = grid.grid_class;
  grid.tree.for_each_active([&](Coord c, T) {
    double xp[3];
    departure_point(grid, vel, c, dt, scheme, tf, xp);
    out.set(c, sample_trilinear(grid, xp));
  });
  return out;
}

// Stage time factors for a step from t to t + dt.
inline void stage_factors(double period, double t, double dt, int scheme, double tf[3]) {
  tf[0] = time_factor(period, t + dt);
  tf[1] = scheme == SCHEME_RK3 ? time_factor(period, t) : 1.0;
  tf[2] = scheme == SCHEME_RK3 ? time_factor(period, t + 0.5 * dt) : 1.0;
}

// ---------------------------------------------------------------- fixtures
// Builder-defined primitives rasterised like src/tools/sphere.rs:17-59:
// the value (float)(sdf) is stored iff |value| < bg, bg = (float)(hw * vs).
struct Primitive {
  int kind = 0;  // 0 sphere, 1 box, 2 noisy sphere
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // sphere: center p0..2, radius p3
  // box:    min p0..2, max p3..5 (exact box distance)
  // noisy:  center p0..2, radius p3, amplitude p4, base frequency p5, seed p6
};

inline double value_noise_lattice(int64_t i, int64_t j, int64_t k, uint64_t seed) {
  uint64_t h = seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)j * 0xC2B2AE3D27D4EB4Full) ^
               ((uint64_t)k * 0x165667B19E3779F9ull);
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // [-1, 1)
}

// Four octaves of trilinear value noise with a cubic fade; octave o has
// frequency f0 * 2^o and weight 2^-o, normalised so |noise| <= amp.
inline double value_noise4(const double w[3], double amp, double f0, uint64_t seed) {
  double sum = 0.0, wsum = 0.0, freq = f0, wt = 1.0;
  for (int o = 0; o < 4; ++o) {
    double g[3], fr[3];
    int64_t b[3];
    for (int a = 0; a < 3; ++a) {
      g[a] = w[a] * freq;
      const double fl = std::floor(g[a]);
      b[a] = (int64_t)fl;
      const double t = g[a] - fl;
      fr[a] = t * t * (3.0 - 2.0 * t);
    }
    double acc = 0.0;
    for (int q = 0; q < 8; ++q) {
      const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
      const double wq = (dx ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) * (dz ? fr[2] : 1.0 - fr[2]);
      acc += wq * value_noise_lattice(b[0] + dx, b[1] + dy, b[2] + dz, seed + (uint64_t)o);
    }
    sum += wt * acc;
    wsum += wt;
    freq *= 2.0;
    wt *= 0.5;
  }
  return amp * (sum / wsum);
}

inline double primitive_sdf(const Primitive& pr, const double w[3]) {
  if (pr.kind == 0 || pr.kind == 2) {
    const double wx = w[0] - pr.p[0], wy = w[1] - pr.p[1], wz = w[2] - pr.p[2];
    double d = std::sqrt(wx * wx + wy * wy + wz * wz) - pr.p[3];
    if (pr.kind == 2) d += value_noise4(w, pr.p[4], pr.p[5], (uint64_t)pr.p[6]);
    return d;
  }
  double q[3], outside = 0.0, inside = -1e300;
  for (int a = 0; a < 3; ++a) {
    const double c = 0.5 * (pr.p[a] + pr.p[a + 3]), hlf = 0.5 * (pr.p[a + 3] - pr.p[a]);
    q[a] = std::fabs(w[a] - c) - hlf;
    const double m = q[a] > 0.0 ? q[a] : 0.0;
    outside += m * m;
    inside = q[a] > inside ? q[a] : inside;
  }
  return std::sqrt(outside) + (inside < 0.0 ? inside : 0.0);
}

// Rasterise one primitive over its padded index box (sphere.rs loop shape).
inline Grid<float> make_primitive_grid(const Primitive& pr, double vs, double hw) {
  const double band = hw * vs;
  const float bg = (float)band;
  Grid<float> g(bg, vs);
  g.grid_class = GC_LEVEL_SET;
  g.name = "sdf";
  g.set_transform(VoxelTransform::uniform(vs));
  double lo[3], hi[3];
  if (pr.kind == 1) {
    for (int a = 0; a < 3; ++a) { lo[a] = pr.p[a] - band; hi[a] = pr.p[a + 3] + band; }
  } else {
    const double ext = pr.p[3] + band + (pr.kind == 2 ? pr.p[4] : 0.0);
    for (int a = 0; a < 3; ++a) { lo[a] = pr.p[a] - ext; hi[a] = pr.p[a] + ext; }
  }
  const double inv = 1.0 / vs;
  int32_t mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = f64_to_i32(std::floor(lo[a] * inv));
    mx[a] = f64_to_i32(std::ceil(hi[a] * inv));
  }
  for (int32_t i = mn[0]; i <= mx[0]; ++i)
    for (int32_t j = mn[1]; j <= mx[1]; ++j)
      for (int32_t k = mn[2]; k <= mx[2]; ++k) {
        const double w[3] = {(double)i * vs, (double)j * vs, (double)k * vs};
        const float sdf = (float)primitive_sdf(pr, w);
        if (std::fabs(sdf) < bg) g.set(Coord(i, j, k), sdf);
      }
  return g;
}

// Left fold of csg_combine over primitives: ((p0 op1 p1) op2 p2) ...
// (src/tools/csg.rs:32-78 applied repeatedly).
inline Grid<float> make_scene_grid(const std::vector<Primitive>& prims, const std::vector<int>& ops, double vs,
                                   double hw) {
  Grid<float> g = make_primitive_grid(prims[0], vs, hw);
  for (size_t i = 1; i < prims.size(); ++i) g = csg_combine(g, make_primitive_grid(prims[i], vs, hw), (CsgOp)ops[i]);
  return g;
}

// ---------------------------------------------------------------- adjoint (FP64)
// Builder-defined: K steps (RK1 or RK3) of a fixed-topology level set
// through a static dense velocity U, loss L = 1/2 sum_active (phi_K - t)^2
// with t = target.get(x).  dL/dU (3 n^3) by reverse-mode differentiation of
// the real-valued step phi_{n+1}(x_i) = S(phi_n, p_i(U)).  Only active voxels
// are variables; inactive and background values are constants.
struct AdjointResult {
  double loss = 0.0;
  std::vector<double> grad[3];
};

// Dense velocity of a double-valued field, for FP64 forward and FD checks.
struct DenseVelocityD {
  int n = 0;
  double h = 1.0;
  std::vector<double> u[3];
  double at(int c, int32_t i, int32_t j, int32_t k) const {
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return 0.0;
    return u[c][((size_t)i * n + j) * n + k];
  }
  void base_frac(const double p[3], int32_t b[3], double f[3]) const {
    const double inv = 1.0 / h;
    for (int a = 0; a < 3; ++a) {
      const double idx = p[a] * inv, fl = std::floor(idx);
      b[a] = f64_to_i32(fl);
      f[a] = idx - fl;
    }
  }
  void sample(const double p[3], double out[3]) const {
    int32_t b[3];
    double f[3];
    base_frac(p, b, f);
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
      for (int q = 0; q < 8; ++q) {
        const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
        const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
        acc += w * at(c, b[0] + dx, b[1] + dy, b[2] + dz);
      }
      out[c] = acc;
    }
  }
  // grad[c][node(p) q] += coef[c] * w_q(p)
  void scatter(const double p[3], const double coef[3], std::vector<double>* grad) const {
    int32_t b[3];
    double f[3];
    base_frac(p, b, f);
    for (int q = 0; q < 8; ++q) {
      const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
      const int32_t i = b[0] + dx, j = b[1] + dy, k = b[2] + dz;
      if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) continue;
      const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
      for (int c = 0; c < 3; ++c) grad[c][((size_t)i * n + j) * n + k] += coef[c] * w;
    }
  }
  // (J^T g)_a = sum_c g_c du_c/dp_a
  void jt(const double p[3], const double g[3], double out[3]) const {
    int32_t b[3];
    double f[3];
    base_frac(p, b, f);
    out[0] = out[1] = out[2] = 0.0;
    for (int q = 0; q < 8; ++q) {
      const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
      const double wx = dx ? f[0] : 1.0 - f[0], wy = dy ? f[1] : 1.0 - f[1], wz = dz ? f[2] : 1.0 - f[2];
      double gu = 0.0;
      for (int c = 0; c < 3; ++c) gu += g[c] * at(c, b[0] + dx, b[1] + dy, b[2] + dz);
      out[0] += gu * (dx ? 1.0 : -1.0) * wy * wz;
      out[1] += gu * wx * (dy ? 1.0 : -1.0) * wz;
      out[2] += gu * wx * wy * (dz ? 1.0 : -1.0);
    }
    for (int a = 0; a < 3; ++a) out[a] /= h;
  }
};

inline double rollout_loss_d(const Grid<double>& phi0, const DenseVelocityD& vel, double dt, int scheme, int steps,
                             const Grid<float>& target) {
  const double tf[3] = {1.0, 1.0, 1.0};
  Grid<double> g = advect_scheme(phi0, vel, dt, scheme, tf);
  for (int s = 1; s < steps; ++s) g = advect_scheme(g, vel, dt, scheme, tf);
  double loss = 0.0;
  g.tree.for_each_active([&](Coord c, double v) {
    const double d = v - (double)target.get(c);
    loss += 0.5 * d * d;
  });
  return loss;
}

inline AdjointResult rollout_adjoint_d(const Grid<double>& phi0, const DenseVelocityD& vel, double dt, int scheme,
                                       int steps, const Grid<float>& target) {
  const double tf[3] = {1.0, 1.0, 1.0};
  std::vector<Grid<double>> states;
  states.push_back(clone_grid(phi0));
  for (int s = 0; s < steps; ++s) states.push_back(advect_scheme(states.back(), vel, dt, scheme, tf));

  AdjointResult res;
  for (int c = 0; c < 3; ++c) res.grad[c].assign((size_t)vel.n * vel.n * vel.n, 0.0);
  std::unordered_map<Coord, double, CoordHash> lam;
  states.back().tree.for_each_active([&](Coord c, double v) {
    const double d = v - (double)target.get(c);
    res.loss += 0.5 * d * d;
    lam[c] = d;
  });
  const double inv_vs = 1.0 / phi0.transform.voxel_size;
  for (int s = steps - 1; s >= 0; --s) {
    const Grid<double>& g = states[s];
    std::unordered_map<Coord, double, CoordHash> lam_prev;
    g.tree.for_each_active([&](Coord c, double) { lam_prev[c] = 0.0; });
    g.tree.for_each_active([&](Coord c, double) {
      const double L = lam.at(c);
      if (L == 0.0) return;
      double x[3], u[3], y1[3], y2[3], p[3];
      g.affine.index_to_world(c, x);
      vel.sample(x, u);
      for (int a = 0; a < 3; ++a) y1[a] = x[a] - dt * u[a];
      if (scheme == SCHEME_RK1) {
        for (int a = 0; a < 3; ++a) p[a] = y1[a];
      } else {
        vel.sample(y1, u);
        for (int a = 0; a < 3; ++a) y2[a] = 0.75 * x[a] + 0.25 * (y1[a] - dt * u[a]);
        vel.sample(y2, u);
        for (int a = 0; a < 3; ++a) p[a] = (1.0 / 3.0) * x[a] + (2.0 / 3.0) * (y2[a] - dt * u[a]);
      }
      double idx[3], f[3];
      int32_t b[3];
      g.transform.world_to_index_f64(p, idx);
      for (int a = 0; a < 3; ++a) {
        const double fl = std::floor(idx[a]);
        b[a] = f64_to_i32(fl);
        f[a] = idx[a] - fl;
      }
      double dS[3] = {0, 0, 0};
      for (int q = 0; q < 8; ++q) {
        const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
        const Coord cc(b[0] + dx, b[1] + dy, b[2] + dz);
        const double cv = g.get(cc);
        const double wx = dx ? f[0] : 1.0 - f[0], wy = dy ? f[1] : 1.0 - f[1], wz = dz ? f[2] : 1.0 - f[2];
        dS[0] += cv * (dx ? 1.0 : -1.0) * wy * wz;
        dS[1] += cv * wx * (dy ? 1.0 : -1.0) * wz;
        dS[2] += cv * wx * wy * (dz ? 1.0 : -1.0);
        if (g.tree.is_active(cc)) lam_prev[cc] += L * wx * wy * wz;
      }
      const double gp[3] = {L * dS[0] * inv_vs, L * dS[1] * inv_vs, L * dS[2] * inv_vs};
      if (scheme == SCHEME_RK1) {
        const double cf[3] = {-dt * gp[0], -dt * gp[1], -dt * gp[2]};
        vel.scatter(x, cf, res.grad);
      } else {
        double jt[3], gy2[3], gy1[3];
        vel.jt(y2, gp, jt);
        for (int a = 0; a < 3; ++a) gy2[a] = (2.0 / 3.0) * gp[a] - (2.0 / 3.0) * dt * jt[a];
        const double c3[3] = {-(2.0 / 3.0) * dt * gp[0], -(2.0 / 3.0) * dt * gp[1], -(2.0 / 3.0) * dt * gp[2]};
        vel.scatter(y2, c3, res.grad);
        vel.jt(y1, gy2, jt);
        for (int a = 0; a < 3; ++a) gy1[a] = 0.25 * gy2[a] - 0.25 * dt * jt[a];
        const double c2[3] = {-0.25 * dt * gy2[0], -0.25 * dt * gy2[1], -0.25 * dt * gy2[2]};
        vel.scatter(y1, c2, res.grad);
        const double c1[3] = {-dt * gy1[0], -dt * gy1[1], -dt * gy1[2]};
        vel.scatter(x, c1, res.grad);
      }
    });
    lam.swap(lam_prev);
  }
  return res;
}

}  // namespace vo
