// volrus oracle: adaptive isosurface meshing — TEST INFRASTRUCTURE ONLY.
//
// Restates, line by line:
//   - src/tools/gradient.rs:13-21     gradient_at
//   - src/tools/gradient.rs:83-135    mean_curvature_at
//   - src/grid.rs:261-277             AffineMap::world_to_index (rounded)
//   - src/tools/adaptive_mesh.rs:47-98, 100-241, 243-271  adaptive_surface
// Only tests/ call these (through oracle_capi.cpp); the product never does.
// f32 arithmetic where the reference uses f32 (build -ffp-contract=off), f64
// elsewhere; acos is the C library's.
#pragma once

#include <unordered_map>

#include "volrus_oracle.hpp"

namespace vo {

// gradient.rs:13-21 (index space)
inline std::array<float, 3> gradient_at(const Grid<float>& g, Coord c) {
  const int32_t x = c.x, y = c.y, z = c.z;
  const float dfdx = (g.get(Coord(x + 1, y, z)) - g.get(Coord(x - 1, y, z))) * 0.5f;
  const float dfdy = (g.get(Coord(x, y + 1, z)) - g.get(Coord(x, y - 1, z))) * 0.5f;
  const float dfdz = (g.get(Coord(x, y, z + 1)) - g.get(Coord(x, y, z - 1))) * 0.5f;
  return {dfdx, dfdy, dfdz};
}

// gradient.rs:56-70
inline float laplacian_at(const Grid<float>& g, Coord c) {
  const int32_t x = c.x, y = c.y, z = c.z;
  const float center = g.get(c);
  const float d2x = g.get(Coord(x + 1, y, z)) - 2.0f * center + g.get(Coord(x - 1, y, z));
  const float d2y = g.get(Coord(x, y + 1, z)) - 2.0f * center + g.get(Coord(x, y - 1, z));
  const float d2z = g.get(Coord(x, y, z + 1)) - 2.0f * center + g.get(Coord(x, y, z - 1));
  return d2x + d2y + d2z;
}

// gradient.rs:83-135
inline float mean_curvature_at(const Grid<float>& g, Coord c) {
  const int32_t x = c.x, y = c.y, z = c.z;
  const float phi_x = (g.get(Coord(x + 1, y, z)) - g.get(Coord(x - 1, y, z))) * 0.5f;
  const float phi_y = (g.get(Coord(x, y + 1, z)) - g.get(Coord(x, y - 1, z))) * 0.5f;
  const float phi_z = (g.get(Coord(x, y, z + 1)) - g.get(Coord(x, y, z - 1))) * 0.5f;
  const float center = g.get(c);
  const float phi_xx = g.get(Coord(x + 1, y, z)) - 2.0f * center + g.get(Coord(x - 1, y, z));
  const float phi_yy = g.get(Coord(x, y + 1, z)) - 2.0f * center + g.get(Coord(x, y - 1, z));
  const float phi_zz = g.get(Coord(x, y, z + 1)) - 2.0f * center + g.get(Coord(x, y, z - 1));
  const float phi_xy = (g.get(Coord(x + 1, y + 1, z)) - g.get(Coord(x + 1, y - 1, z)) - g.get(Coord(x - 1, y + 1, z)) +
                        g.get(Coord(x - 1, y - 1, z))) *
                       0.25f;
  const float phi_xz = (g.get(Coord(x + 1, y, z + 1)) - g.get(Coord(x + 1, y, z - 1)) - g.get(Coord(x - 1, y, z + 1)) +
                        g.get(Coord(x - 1, y, z - 1))) *
                       0.25f;
  const float phi_yz = (g.get(Coord(x, y + 1, z + 1)) - g.get(Coord(x, y + 1, z - 1)) - g.get(Coord(x, y - 1, z + 1)) +
                        g.get(Coord(x, y - 1, z - 1))) *
                       0.25f;
  const float grad_mag_sq = phi_x * phi_x + phi_y * phi_y + phi_z * phi_z;
  const float grad_mag = std::sqrt(grad_mag_sq);
  if (grad_mag < 1e-10f) return 0.0f;
  const float numerator = phi_xx * (phi_y * phi_y + phi_z * phi_z) + phi_yy * (phi_x * phi_x + phi_z * phi_z) +
                          phi_zz * (phi_x * phi_x + phi_y * phi_y) -
                          2.0f * (phi_x * phi_y * phi_xy + phi_x * phi_z * phi_xz + phi_y * phi_z * phi_yz);
  return numerator / (grad_mag_sq * grad_mag);
}

// grid.rs:261-277: inverse affine, then f64::round (half away from zero) and `as i32`.
inline Coord world_to_index(const AffineMap& m, const double w[3]) {
  double f[3];
  for (int r = 0; r < 3; ++r) f[r] = m.inv[r][0] * w[0] + m.inv[r][1] * w[1] + m.inv[r][2] * w[2] + m.inv[r][3];
  return Coord(f64_to_i32(std::round(f[0])), f64_to_i32(std::round(f[1])), f64_to_i32(std::round(f[2])));
}

// adaptive_mesh.rs:47-55
inline double curvature_at_world(const Grid<float>& g, const double w[3]) {
  const Coord c = world_to_index(g.affine, w);
  const float h = mean_curvature_at(g, c);
  return (double)h / g.affine.voxel_size;
}

// adaptive_mesh.rs:58-67 (the squared length is summed in f32, then widened)
inline std::array<double, 3> normal_at_world(const Grid<float>& g, const double w[3]) {
  const Coord c = world_to_index(g.affine, w);
  const std::array<float, 3> gr = gradient_at(g, c);
  const double len = std::sqrt((double)(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]));
  if (len < 1e-15) return {0.0, 0.0, 1.0};
  return {(double)gr[0] / len, (double)gr[1] / len, (double)gr[2] / len};
}

// adaptive_mesh.rs:70-78
inline std::array<double, 3> project_onto_isosurface(const Grid<float>& g, const std::array<double, 3>& p, float iso) {
  const float sdf = sample_trilinear(g, p.data()) - iso;
  const std::array<double, 3> n = normal_at_world(g, p.data());
  return {p[0] - (double)sdf * n[0], p[1] - (double)sdf * n[1], p[2] - (double)sdf * n[2]};
}

// adaptive_mesh.rs:80-94
inline double edge_length(const std::array<double, 3>& a, const std::array<double, 3>& b) {
  const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}
inline std::array<double, 3> midpoint(const std::array<double, 3>& a, const std::array<double, 3>& b) {
  return {(a[0] + b[0]) * 0.5, (a[1] + b[1]) * 0.5, (a[2] + b[2]) * 0.5};
}

// adaptive_mesh.rs:261-271
inline double angle_between(const std::array<double, 3>& a, const std::array<double, 3>& b) {
  const double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (la < 1e-15 || lb < 1e-15) return 0.0;
  double c = dot / (la * lb);
  c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);  // f64::clamp (NaN stays NaN)
  return std::acos(c);
}

// f64::max: the other operand when one is NaN.
inline double rmax(double a, double b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }

struct RenderMesh {
  std::vector<std::array<double, 3>> vertices, normals;
  std::vector<std::array<uint32_t, 3>> triangles;
  std::vector<double> curvatures;
};

struct PairHashU32 {
  size_t operator()(const std::pair<uint32_t, uint32_t>& p) const {
    return std::hash<uint64_t>()(((uint64_t)p.first << 32) | p.second);
  }
};

// adaptive_mesh.rs:100-241
inline RenderMesh adaptive_surface(const Grid<float>& g, float isovalue, double max_edge_length,
                                   double curvature_threshold, uint32_t max_depth) {
  const Mesh base = volume_to_mesh(g, isovalue);
  std::vector<std::array<double, 3>> vertices = base.vertices;
  std::vector<std::array<uint32_t, 3>> triangles = base.triangles;
  for (uint32_t depth = 0; depth < max_depth; ++depth) {
    const double min_edge = max_edge_length / (double)(1ull << (depth + 1));
    std::vector<double> vc(vertices.size());
    for (size_t i = 0; i < vertices.size(); ++i) vc[i] = std::fabs(curvature_at_world(g, vertices[i].data()));
    std::unordered_map<std::pair<uint32_t, uint32_t>, uint32_t, PairHashU32> cache;
    std::vector<std::array<uint32_t, 3>> nt;
    auto mid = [&](uint32_t a, uint32_t b) {  // :243-259
      const std::pair<uint32_t, uint32_t> key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto it = cache.find(key);
      if (it != cache.end()) return it->second;
      const std::array<double, 3> m = midpoint(vertices[a], vertices[b]);
      const std::array<double, 3> p = project_onto_isosurface(g, m, isovalue);
      const uint32_t idx = (uint32_t)vertices.size();
      vertices.push_back(p);
      cache.emplace(key, idx);
      return idx;
    };
    for (const auto& t : triangles) {
      const double max_curv = rmax(rmax(vc[t[0]], vc[t[1]]), vc[t[2]]);
      const double e01 = edge_length(vertices[t[0]], vertices[t[1]]);
      const double e12 = edge_length(vertices[t[1]], vertices[t[2]]);
      const double e20 = edge_length(vertices[t[2]], vertices[t[0]]);
      const double max_edge = rmax(rmax(e01, e12), e20);
      if (max_curv > curvature_threshold && max_edge > min_edge) {
        const uint32_t m01 = mid(t[0], t[1]);
        const uint32_t m12 = mid(t[1], t[2]);
        const uint32_t m20 = mid(t[2], t[0]);
        nt.push_back({t[0], m01, m20});
        nt.push_back({t[1], m12, m01});
        nt.push_back({t[2], m20, m12});
        nt.push_back({m01, m12, m20});
      } else {
        nt.push_back(t);
      }
    }
    triangles = std::move(nt);
  }
  RenderMesh out;
  out.normals.assign(vertices.size(), {0.0, 0.0, 0.0});
  for (const auto& t : triangles) {  // :166-211
    const auto& v0 = vertices[t[0]];
    const auto& v1 = vertices[t[1]];
    const auto& v2 = vertices[t[2]];
    const std::array<double, 3> e01 = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const std::array<double, 3> e02 = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const std::array<double, 3> e12 = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
    const double fx = e01[1] * e02[2] - e01[2] * e02[1];
    const double fy = e01[2] * e02[0] - e01[0] * e02[2];
    const double fz = e01[0] * e02[1] - e01[1] * e02[0];
    const double a0 = angle_between(e01, e02);
    const double a1 = angle_between({-e01[0], -e01[1], -e01[2]}, e12);
    const double a2 = angle_between({-e02[0], -e02[1], -e02[2]}, {-e12[0], -e12[1], -e12[2]});
    const double ang[3] = {a0, a1, a2};
    for (int k = 0; k < 3; ++k) {
      auto& n = out.normals[t[k]];
      n[0] += fx * ang[k];
      n[1] += fy * ang[k];
      n[2] += fz * ang[k];
    }
  }
  for (auto& n : out.normals) {  // :214-222
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 1e-15) {
      n[0] /= len;
      n[1] /= len;
      n[2] /= len;
    }
  }
  out.curvatures.resize(vertices.size());  // :225-229
  for (size_t i = 0; i < vertices.size(); ++i) out.curvatures[i] = curvature_at_world(g, vertices[i].data());
  out.vertices = std::move(vertices);
  out.triangles = std::move(triangles);
  return out;
}

}  // namespace vo
