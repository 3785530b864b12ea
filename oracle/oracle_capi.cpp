// C entry points over the oracle for ctypes — TEST INFRASTRUCTURE ONLY.
//
// Loaded by tests/, by the smoke() check and by bench.py's cpu_baseline leg;
// never by the product library.  Grids cross this boundary as opaque handles
// and as NanoGrid bytes (src/nano/layout.rs:13-52 layout).
#include <chrono>
#include <cstdio>

#include "volrus_oracle_ext.hpp"
#include "volrus_oracle_io.hpp"
#include "volrus_oracle_adaptive.hpp"

using namespace vo;

namespace {
struct OMesh {
  Mesh m;
};
Grid<float>* G(void* p) { return static_cast<Grid<float>*>(p); }
Grid<double>* GD(void* p) { return static_cast<Grid<double>*>(p); }

DenseVelocity make_dense(int n, double h, const float* u) {
  DenseVelocity d;
  d.n = n;
  d.h = h;
  const size_t nn = (size_t)n * n * n;
  for (int c = 0; c < 3; ++c) d.u[c].assign(u + c * nn, u + (c + 1) * nn);
  return d;
}
AnalyticVelocity make_analytic(int kind, const double* a6) {
  AnalyticVelocity v;
  v.kind = kind;
  for (int i = 0; i < 6; ++i) v.a[i] = a6[i];
  return v;
}
std::vector<Primitive> make_prims(int n, const int* kinds, const double* params) {
  std::vector<Primitive> p((size_t)n);
  for (int i = 0; i < n; ++i) {
    p[i].kind = kinds[i];
    for (int j = 0; j < 8; ++j) p[i].p[j] = params[i * 8 + j];
  }
  return p;
}
}  // namespace

extern "C" {

// ---------------------------------------------------------------- grids
void* or_new_grid(float bg, double vs) { return new Grid<float>(bg, vs); }
void* or_new_level_set(double vs, double hw) { return new Grid<float>(Grid<float>::level_set(vs, hw)); }
void or_free(void* g) { delete G(g); }
void or_free_f64(void* g) { delete GD(g); }
void or_set(void* g, int x, int y, int z, float v) { G(g)->set(Coord(x, y, z), v); }
float or_get(void* g, int x, int y, int z) { return G(g)->get(Coord(x, y, z)); }
int or_is_active(void* g, int x, int y, int z) { return G(g)->tree.is_active(Coord(x, y, z)) ? 1 : 0; }
void or_deactivate(void* g, int x, int y, int z) { G(g)->tree.deactivate(Coord(x, y, z)); }
void or_set_class(void* g, int c) { G(g)->grid_class = (GridClass)c; }
void or_set_transform_uniform(void* g, double vs) { G(g)->set_transform(VoxelTransform::uniform(vs)); }
uint64_t or_active_count(void* g) { return G(g)->active_voxel_count(); }
uint64_t or_leaf_count(void* g) { return G(g)->leaf_count(); }
float or_background(void* g) { return G(g)->background(); }
double or_sample_trilinear(void* g, double x, double y, double z) {
  const double w[3] = {x, y, z};
  return sample_trilinear(*G(g), w);
}

// ---------------------------------------------------------------- nano
size_t or_to_nano(void* g, uint8_t* out, size_t cap) {
  const std::vector<uint8_t> b = nano_from_grid(*G(g));
  if (out && cap >= b.size()) std::memcpy(out, b.data(), b.size());
  return b.size();
}
int or_nano_validate(const uint8_t* b, size_t len, char* msg, size_t cap) {
  const std::string e = nano_validate(b, len);
  if (msg && cap) std::snprintf(msg, cap, "%s", e.c_str());
  return e.empty() ? 0 : 1;
}
void* or_from_nano(const uint8_t* b, size_t len) {
  if (!nano_validate(b, len).empty()) return nullptr;
  return new Grid<float>(nano_to_grid(b));
}
float or_nano_get(const uint8_t* b, int x, int y, int z) { return nano_get(b, Coord(x, y, z)); }
int or_nano_is_active(const uint8_t* b, int x, int y, int z) { return nano_is_active(b, Coord(x, y, z)) ? 1 : 0; }

// ---------------------------------------------------------------- fixtures
void* or_sphere(double r, double cx, double cy, double cz, double vs, double hw) {
  const double c[3] = {cx, cy, cz};
  return new Grid<float>(make_level_set_sphere<float>(r, c, vs, hw));
}
void* or_csg(void* a, void* b, int op) { return new Grid<float>(csg_combine(*G(a), *G(b), (CsgOp)op)); }
void* or_scene(int n, const int* kinds, const double* params, const int* ops, double vs, double hw) {
  const std::vector<Primitive> p = make_prims(n, kinds, params);
  const std::vector<int> o(ops, ops + n);
  return new Grid<float>(make_scene_grid(p, o, vs, hw));
}
double or_primitive_sdf(int kind, const double* params, double x, double y, double z) {
  Primitive p;
  p.kind = kind;
  for (int j = 0; j < 8; ++j) p.p[j] = params[j];
  const double w[3] = {x, y, z};
  return primitive_sdf(p, w);
}
void or_sincospi(double x, double* s, double* c) { sincospi_det(x, s, c); }
uint64_t or_splitmix_next(uint64_t* state) {
  SplitMix64 r(*state);
  const uint64_t v = r.next();
  *state = r.s;
  return v;
}
double or_value_noise4(double x, double y, double z, double amp, double f0, uint64_t seed) {
  const double w[3] = {x, y, z};
  return value_noise4(w, amp, f0, seed);
}
double or_value_noise_lattice(int64_t i, int64_t j, int64_t k, uint64_t seed) {
  return value_noise_lattice(i, j, k, seed);
}

// ---------------------------------------------------------------- advection
// Reference RK1 step with a VelocityField of three grids (advection.rs:32-56).
void* or_advect_grids3(void* g, void* vx, void* vy, void* vz, double dt) {
  VelocityField<float> vf{clone_grid(*G(vx)), clone_grid(*G(vy)), clone_grid(*G(vz))};
  return new Grid<float>(advect_level_set(*G(g), vf, dt));
}
// Builder-defined schemes; tf = stage time factors.
void* or_advect_grids3_scheme(void* g, void* vx, void* vy, void* vz, double dt, int scheme, const double* tf) {
  VelocityField<float> vf{clone_grid(*G(vx)), clone_grid(*G(vy)), clone_grid(*G(vz))};
  return new Grid<float>(advect_scheme(*G(g), vf, dt, scheme, tf));
}
void* or_advect_analytic(void* g, int kind, const double* a6, double dt, int scheme, const double* tf) {
  return new Grid<float>(advect_scheme(*G(g), make_analytic(kind, a6), dt, scheme, tf));
}
void* or_advect_dense(void* g, int n, double h, const float* u, double dt, int scheme, const double* tf) {
  return new Grid<float>(advect_scheme(*G(g), make_dense(n, h, u), dt, scheme, tf));
}
void or_stage_factors(double period, double t, double dt, int scheme, double* tf) {
  stage_factors(period, t, dt, scheme, tf);
}

// FP64 forms: the same step with double values (input converted from float).
void* or_to_f64(void* g) { return new Grid<double>(convert_grid<double, float>(*G(g))); }
void* or_advect_analytic_f64(void* g, int kind, const double* a6, double dt, int scheme, const double* tf) {
  return new Grid<double>(advect_scheme(*GD(g), make_analytic(kind, a6), dt, scheme, tf));
}
void* or_advect_dense_f64(void* g, int n, double h, const float* u, double dt, int scheme, const double* tf) {
  return new Grid<double>(advect_scheme(*GD(g), make_dense(n, h, u), dt, scheme, tf));
}
double or_get_f64(void* g, int x, int y, int z) { return GD(g)->get(Coord(x, y, z)); }

// ---------------------------------------------------------------- band tools
size_t or_zero_crossings(void* g, int32_t* out, size_t cap) {
  const std::vector<Coord> c = find_zero_crossings(*G(g));
  for (size_t i = 0; i < c.size() && i < cap; ++i) {
    out[3 * i] = c[i].x;
    out[3 * i + 1] = c[i].y;
    out[3 * i + 2] = c[i].z;
  }
  return c.size();
}
void or_trim(void* g, float hw) { trim_narrow_band(*G(g), hw); }
void or_rebuild(void* g, float hw) { rebuild_narrow_band(*G(g), hw); }
void or_rebuild_bounded(void* g, float hw) { rebuild_narrow_band_bounded(*G(g), hw); }

// ---------------------------------------------------------------- marching cubes
void* or_mesh(void* g, float iso) { return new OMesh{volume_to_mesh(*G(g), iso)}; }
size_t or_mesh_nv(void* m) { return static_cast<OMesh*>(m)->m.vertices.size(); }
size_t or_mesh_nt(void* m) { return static_cast<OMesh*>(m)->m.triangles.size(); }
void or_mesh_copy(void* m, double* v, uint32_t* t) {
  const Mesh& mm = static_cast<OMesh*>(m)->m;
  for (size_t i = 0; i < mm.vertices.size(); ++i)
    for (int a = 0; a < 3; ++a) v[3 * i + a] = mm.vertices[i][a];
  for (size_t i = 0; i < mm.triangles.size(); ++i)
    for (int a = 0; a < 3; ++a) t[3 * i + a] = mm.triangles[i][a];
}
int or_mesh_watertight(void* m) { return static_cast<OMesh*>(m)->m.is_watertight() ? 1 : 0; }
void or_mesh_free(void* m) { delete static_cast<OMesh*>(m); }
size_t or_mc_tiles(void* g, int32_t* out, size_t cap) {
  const std::vector<Coord> t = mc_tiles(*G(g));
  for (size_t i = 0; i < t.size() && i < cap; ++i) {
    out[3 * i] = t[i].x;
    out[3 * i + 1] = t[i].y;
    out[3 * i + 2] = t[i].z;
  }
  return t.size();
}

// ---------------------------------------------------------------- adjoint (FP64)
static DenseVelocityD make_dense_d(int n, double h, const double* u) {
  DenseVelocityD d;
  d.n = n;
  d.h = h;
  const size_t nn = (size_t)n * n * n;
  for (int c = 0; c < 3; ++c) d.u[c].assign(u + c * nn, u + (c + 1) * nn);
  return d;
}
double or_rollout_loss_f64(void* g, int n, double h, const double* u, double dt, int scheme, int steps, void* target) {
  const Grid<double> gd = convert_grid<double, float>(*G(g));
  return rollout_loss_d(gd, make_dense_d(n, h, u), dt, scheme, steps, *G(target));
}
double or_rollout_adjoint_f64(void* g, int n, double h, const double* u, double dt, int scheme, int steps,
                              void* target, double* grad) {
  const Grid<double> gd = convert_grid<double, float>(*G(g));
  const AdjointResult r = rollout_adjoint_d(gd, make_dense_d(n, h, u), dt, scheme, steps, *G(target));
  const size_t nn = (size_t)n * n * n;
  for (int c = 0; c < 3; ++c) std::memcpy(grad + c * nn, r.grad[c].data(), nn * sizeof(double));
  return r.loss;
}
// f32 forward rollout (the GPU's arithmetic) and its loss against a target grid.
void* or_rollout_dense(void* g, int n, double h, const float* u, double dt, int scheme, int steps) {
  const double tf[3] = {1.0, 1.0, 1.0};
  const DenseVelocity d = make_dense(n, h, u);
  Grid<float> cur = advect_scheme(*G(g), d, dt, scheme, tf);
  for (int s = 1; s < steps; ++s) cur = advect_scheme(cur, d, dt, scheme, tf);
  return new Grid<float>(std::move(cur));
}

// ---------------------------------------------------------------- CPU timing
// Wall time of one reference RK1 advect step plus marching cubes on `g` with
// an analytic field (used for the reported CPU baseline).  Returns seconds.
double or_time_advect_mc(void* g, int kind, const double* a6, double dt, float iso, uint64_t* nvert) {
  const double tf[3] = {1.0, 1.0, 1.0};
  const auto t0 = std::chrono::steady_clock::now();
  Grid<float> out = advect_scheme(*G(g), make_analytic(kind, a6), dt, SCHEME_RK1, tf);
  Mesh m = volume_to_mesh(out, iso);
  const auto t1 = std::chrono::steady_clock::now();
  if (nvert) *nvert = m.vertices.size();
  return std::chrono::duration<double>(t1 - t0).count();
}
double or_time_step(void* g, int kind, const double* a6, double dt, int scheme, const double* tf, int do_rebuild,
                    float hw, int do_mc, float iso) {
  const auto t0 = std::chrono::steady_clock::now();
  Grid<float> out = advect_scheme(*G(g), make_analytic(kind, a6), dt, scheme, tf);
  if (do_rebuild) rebuild_narrow_band_bounded(out, hw);
  if (do_mc) {
    Mesh m = volume_to_mesh(out, iso);
    (void)m;
  }
  const auto t1 = std::chrono::steady_clock::now();
  *G(g) = std::move(out);
  return std::chrono::duration<double>(t1 - t0).count();
}

// ---------------------------------------------------------------- .vol, velocity extension
size_t or_vol_write(void* g, uint8_t* out, size_t cap) {
  const std::vector<uint8_t> b = vol_write(*G(g));
  if (out && cap >= b.size()) std::memcpy(out, b.data(), b.size());
  return b.size();
}
// The grid read, or null with the error text in msg.
void* or_vol_read(const uint8_t* b, size_t len, char* msg, size_t cap) {
  Grid<float> g(0.0f, 1.0);
  const std::string e = vol_read(b, len, &g);
  if (msg && cap) std::snprintf(msg, cap, "%s", e.c_str());
  return e.empty() ? new Grid<float>(std::move(g)) : nullptr;
}
void or_extend_velocity(void* sdf, void* vx, void* vy, void* vz, float half_width) {
  extend_velocity(*G(sdf), *G(vx), *G(vy), *G(vz), half_width);
}

// ---------------------------------------------------------------- differential operators, adaptive meshing
void or_gradient_at(void* g, int x, int y, int z, float* out) {
  const std::array<float, 3> v = gradient_at(*G(g), Coord(x, y, z));
  out[0] = v[0];
  out[1] = v[1];
  out[2] = v[2];
}
float or_laplacian_at(void* g, int x, int y, int z) { return laplacian_at(*G(g), Coord(x, y, z)); }
float or_mean_curvature_at(void* g, int x, int y, int z) { return mean_curvature_at(*G(g), Coord(x, y, z)); }
void* or_adaptive_surface(void* g, float iso, double max_edge, double threshold, uint32_t depth) {
  return new RenderMesh(adaptive_surface(*G(g), iso, max_edge, threshold, depth));
}
size_t or_render_nv(void* m) { return static_cast<RenderMesh*>(m)->vertices.size(); }
size_t or_render_nt(void* m) { return static_cast<RenderMesh*>(m)->triangles.size(); }
void or_render_copy(void* m, double* v, double* n, uint32_t* t, double* c) {
  const RenderMesh& r = *static_cast<RenderMesh*>(m);
  for (size_t i = 0; i < r.vertices.size(); ++i)
    for (int k = 0; k < 3; ++k) {
      v[3 * i + k] = r.vertices[i][k];
      n[3 * i + k] = r.normals[i][k];
    }
  for (size_t i = 0; i < r.triangles.size(); ++i)
    for (int k = 0; k < 3; ++k) t[3 * i + k] = r.triangles[i][k];
  for (size_t i = 0; i < r.curvatures.size(); ++i) c[i] = r.curvatures[i];
}
void or_render_free(void* m) { delete static_cast<RenderMesh*>(m); }

}  // extern "C"
