// volrus oracle: .vol I/O and velocity extension — TEST INFRASTRUCTURE ONLY.
//
// Restates, line by line:
//   - src/io/vol_format.rs:53-85   write_vol
//   - src/io/vol_format.rs:88-147  read_vol
//   - src/tools/level_set_track.rs:127-175  extend_velocity (+ dist_sq)
// Only tests/ call these (through oracle_capi.cpp); the product never does.
#pragma once

#include <cstdio>

#include "volrus_oracle.hpp"

namespace vo {

// The reference iterates active voxels root node by root node in HashMap
// order, then child slot by child slot (src/tree/root.rs:104-106,
// internal.rs:319-330), then by offset.  With one root node that order is
// fixed; across root nodes it is not, and the oracle visits root nodes in
// ascending origin order (builder-defined; equal to the reference whenever a
// grid lies in one 128^3 node).
template <class T, class F>
void for_each_active_root_sorted(const Grid<T>& g, F&& f) {
  std::vector<Coord> keys;
  for (const auto& kv : g.tree.tiles) keys.push_back(kv.first);
  std::sort(keys.begin(), keys.end());
  for (const Coord& k : keys) {
    const InternalNode<T>& node = g.tree.tiles.at(k);
    for (size_t i = 0; i < (size_t)INTERNAL_SIZE; ++i)
      if (node.has_child(i)) node.children[i]->for_each_active(f);
  }
}

// ---------------------------------------------------------------- .vol
constexpr size_t VOL_HEADER = 32;
constexpr size_t VOL_LEAF = 12 + 64 + 4 * LEAF_SIZE;  // origin, mask, values

inline void vol_put(std::vector<uint8_t>& b, const void* p, size_t n) {
  const uint8_t* q = static_cast<const uint8_t*>(p);
  b.insert(b.end(), q, q + n);
}

// vol_format.rs:53-85.  Leaves in ascending origin order (the reference
// writes them in its HashMap order; any order reads back the same grid).
inline std::vector<uint8_t> vol_write(const Grid<float>& g) {
  std::vector<uint8_t> b;
  vol_put(b, "VOLR", 4);
  const uint32_t version = 1, cls = (uint32_t)g.grid_class, n = (uint32_t)g.leaf_count();
  vol_put(b, &version, 4);
  vol_put(b, &cls, 4);
  vol_put(b, &g.transform.voxel_size, 8);
  const float bg = g.background();
  vol_put(b, &bg, 4);
  vol_put(b, &n, 4);
  const uint8_t reserved[4] = {0, 0, 0, 0};
  vol_put(b, reserved, 4);
  std::vector<const LeafNode<float>*> leaves;
  g.tree.for_each_leaf([&](const LeafNode<float>& l) { leaves.push_back(&l); });
  std::sort(leaves.begin(), leaves.end(),
            [](const LeafNode<float>* a, const LeafNode<float>* c) { return a->origin < c->origin; });
  for (const LeafNode<float>* l : leaves) {
    vol_put(b, &l->origin.x, 4);
    vol_put(b, &l->origin.y, 4);
    vol_put(b, &l->origin.z, 4);
    vol_put(b, l->active, 64);
    vol_put(b, l->values.data(), 4 * LEAF_SIZE);
  }
  return b;
}

// vol_format.rs:88-147.  Returns the error text ("" on success) and fills
// *out.  read_exact's short-read error reads "failed to fill whole buffer".
inline std::string vol_read(const uint8_t* b, size_t len, Grid<float>* out) {
  if (len < VOL_HEADER) return "failed to fill whole buffer";
  if (std::memcmp(b, "VOLR", 4) != 0) return "not a .vol file (bad magic)";
  uint32_t version, cls, n;
  double vs;
  float bg;
  std::memcpy(&version, b + 4, 4);
  if (version != 1) return "unsupported .vol version: " + std::to_string(version);
  std::memcpy(&cls, b + 8, 4);
  if (cls > 2) return "unknown grid class: " + std::to_string(cls);
  std::memcpy(&vs, b + 12, 8);
  std::memcpy(&bg, b + 20, 4);
  std::memcpy(&n, b + 24, 4);
  Grid<float> g(bg, vs);
  g.grid_class = (GridClass)cls;
  size_t at = VOL_HEADER;
  for (uint32_t i = 0; i < n; ++i) {
    if (len - at < VOL_LEAF) return "failed to fill whole buffer";
    int32_t o[3];
    std::memcpy(o, b + at, 12);
    const Coord origin(o[0], o[1], o[2]);
    uint64_t mask[8];
    std::memcpy(mask, b + at + 12, 64);
    std::array<float, LEAF_SIZE> values;
    std::memcpy(values.data(), b + at + 76, 4 * LEAF_SIZE);
    at += VOL_LEAF;
    g.set(origin, values[0]);  // :128-130: allocates the leaf (origin & ~7)
    LeafNode<float>* leaf = const_cast<LeafNode<float>*>(g.tree.leaf_at(origin));  // :132-137
    leaf->values = values;
    std::memcpy(leaf->active, mask, 64);
  }
  *out = std::move(g);
  return "";
}

// ---------------------------------------------------------------- extend_velocity
// level_set_track.rs:171-175
inline int64_t dist_sq(Coord a, Coord b) {
  const int64_t dx = (int64_t)a.x - b.x, dy = (int64_t)a.y - b.y, dz = (int64_t)a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

// level_set_track.rs:127-169.  Interface voxels: |v| < vs (f32); band voxels:
// vs <= |v| < half_width * vs; each band voxel takes the velocity of the
// first interface voxel (in iteration order) at minimum squared distance
// (min_by_key keeps the first minimum), set() activating it in vx, vy, vz.
inline void extend_velocity(const Grid<float>& sdf, Grid<float>& vx, Grid<float>& vy, Grid<float>& vz,
                            float half_width) {
  const float voxel_size = (float)sdf.transform.voxel_size;
  const float band = half_width * voxel_size;
  std::vector<Coord> iface, bvox;
  for_each_active_root_sorted(sdf, [&](Coord c, float v) {
    if (std::fabs(v) < voxel_size) iface.push_back(c);
  });
  if (iface.empty()) return;
  for_each_active_root_sorted(sdf, [&](Coord c, float v) {
    if (std::fabs(v) >= voxel_size && std::fabs(v) < band) bvox.push_back(c);
  });
  for (const Coord& c : bvox) {
    size_t best = 0;
    int64_t bd = dist_sq(c, iface[0]);
    for (size_t k = 1; k < iface.size(); ++k) {
      const int64_t d = dist_sq(c, iface[k]);
      if (d < bd) {
        bd = d;
        best = k;
      }
    }
    const Coord s = iface[best];
    vx.set(c, vx.get(s));
    vy.set(c, vy.get(s));
    vz.set(c, vz.get(s));
  }
}

}  // namespace vo
