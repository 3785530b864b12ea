// volrus oracle — CPU restatement of the reference's narrow-band level-set path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
// entry point and bench.py's cpu_baseline leg may build and call it, and
// nothing else.  The product library (sharc.volrus_amd/) never includes,
// links or calls anything under oracle/.
//
// It restates, line by line, the Rust reference (revred/Sharc.Volrus):
//   - tree:        src/math/coord.rs:8-83, src/tree/leaf.rs:18-186,
//                  src/tree/internal.rs:39-247, src/tree/root.rs:16-171
//   - grid:        src/grid.rs:62-122, 131-296, 349-514
//   - trilinear:   src/tools/interpolation.rs:11-48
//   - advection:   src/tools/advection.rs:10-56
//   - band:        src/tools/level_set_track.rs:14-114
//   - MC:          src/tools/volume_to_mesh.rs:11-201
//   - Nano:        src/nano/layout.rs:7-285, src/nano/lookup.rs:11-79
//   - fixtures:    src/tools/sphere.rs:17-59, src/tools/csg.rs:15-78
// and adds the parts the reference does not have, each marked
// "builder-defined": RK3 back-tracing, analytic and dense velocity fields,
// the bounded-radius rebuild, the adjoint of the RK1 step, and fixture
// generators (slot box, noisy sphere, sphere unions, SplitMix64).
//
// Everything is templated on the value type T: T = float restates the
// reference bit for bit (build with -ffp-contract=off, no fast-math, SSE2),
// T = double gives the FP64 comparison used to state errors.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../sharc.volrus_amd/csrc/mc_tables.h"

namespace vo {

// ---------------------------------------------------------------- Coord
// src/math/coord.rs:8-13; derived Ord is lexicographic (x, y, z).
struct Coord {
  int32_t x = 0, y = 0, z = 0;
  Coord() = default;
  Coord(int32_t x_, int32_t y_, int32_t z_) : x(x_), y(y_), z(z_) {}
  bool operator==(const Coord& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Coord& o) const { return !(*this == o); }
  bool operator<(const Coord& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    return z < o.z;
  }
  Coord operator+(const Coord& o) const { return Coord(x + o.x, y + o.y, z + o.z); }
};

struct CoordHash {
  size_t operator()(const Coord& c) const {
    uint64_t h = (uint64_t)(uint32_t)c.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)c.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)c.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// src/math/coord.rs:63-70
inline Coord aligned(Coord c, unsigned log2dim) {
  const int32_t mask = ~((int32_t(1) << log2dim) - 1);
  return Coord(c.x & mask, c.y & mask, c.z & mask);
}

// src/math/coord.rs:74-83 (ZYX order: x * 64 + y * 8 + z for leaves)
inline size_t offset_in_tile(Coord c, unsigned log2dim) {
  const int32_t mask = (int32_t(1) << log2dim) - 1;
  const size_t lx = (size_t)(c.x & mask), ly = (size_t)(c.y & mask), lz = (size_t)(c.z & mask);
  return (lx << (2 * log2dim)) | (ly << log2dim) | lz;
}

// Rust `f64 as i32`: saturating, NaN -> 0.
inline int32_t f64_to_i32(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
  return (int32_t)v;
}

constexpr unsigned LEAF_LOG2DIM = 3;           // src/tree/leaf.rs:8
constexpr int LEAF_SIZE = 512;                 // src/tree/leaf.rs:10
constexpr int32_t LEAF_DIM = 8;                // src/tree/leaf.rs:12
constexpr unsigned INTERNAL_LOG2DIM = 4;       // src/tree/internal.rs:14
constexpr int INTERNAL_SIZE = 4096;            // src/tree/internal.rs:18
constexpr unsigned INTERNAL_TOTAL_LOG2DIM = 7; // src/tree/internal.rs:23

// ---------------------------------------------------------------- LeafNode
// src/tree/leaf.rs:18-186
template <class T>
struct LeafNode {
  Coord origin;
  std::array<T, LEAF_SIZE> values;
  uint64_t active[8];

  LeafNode(Coord o, T background) : origin(aligned(o, LEAF_LOG2DIM)) {  // :29-35
    values.fill(background);
    std::memset(active, 0, sizeof(active));
  }
  T get(Coord c) const { return values[offset_in_tile(c, LEAF_LOG2DIM)]; }  // :45-48
  void set(Coord c, T v) {                                                 // :52-56
    const size_t off = offset_in_tile(c, LEAF_LOG2DIM);
    values[off] = v;
    active[off / 64] |= 1ull << (off % 64);
  }
  void set_inactive(Coord c, T background) {  // :60-64
    const size_t off = offset_in_tile(c, LEAF_LOG2DIM);
    values[off] = background;
    active[off / 64] &= ~(1ull << (off % 64));
  }
  bool is_active(Coord c) const {  // :68-71
    const size_t off = offset_in_tile(c, LEAF_LOG2DIM);
    return ((active[off / 64] >> (off % 64)) & 1ull) == 1ull;
  }
  uint32_t active_count() const {
    uint32_t n = 0;
    for (int w = 0; w < 8; ++w) n += (uint32_t)__builtin_popcountll(active[w]);
    return n;
  }
  bool is_empty() const {  // :86-88
    for (int w = 0; w < 8; ++w)
      if (active[w]) return false;
    return true;
  }
  // :145-165, bit order within each word (trailing zeros first).
  template <class F>
  void for_each_active(F&& f) const {
    for (int w = 0; w < 8; ++w) {
      uint64_t word = active[w];
      while (word) {
        const int bit = __builtin_ctzll(word);
        word &= word - 1;
        const int off = w * 64 + bit;
        f(Coord(origin.x + (off >> 6), origin.y + ((off >> 3) & 7), origin.z + (off & 7)), values[off]);
      }
    }
  }
};

// ---------------------------------------------------------------- InternalNode
// src/tree/internal.rs:39-247.  Slots are either a child leaf or a tile.  The
// path never calls prune(), so every tile holds the background value and no
// tile is active; that is what this restatement stores.
template <class T>
struct InternalNode {
  Coord origin;
  std::vector<std::unique_ptr<LeafNode<T>>> children;
  uint64_t child_mask[64];
  T background;

  InternalNode(Coord o, T bg) : origin(aligned(o, INTERNAL_TOTAL_LOG2DIM)), children(INTERNAL_SIZE), background(bg) {
    std::memset(child_mask, 0, sizeof(child_mask));
  }
  InternalNode(InternalNode&&) = default;

  size_t child_index(Coord c) const {  // :88-95
    const size_t lx = (size_t)((c.x - origin.x) >> LEAF_LOG2DIM);
    const size_t ly = (size_t)((c.y - origin.y) >> LEAF_LOG2DIM);
    const size_t lz = (size_t)((c.z - origin.z) >> LEAF_LOG2DIM);
    return (lx << (2 * INTERNAL_LOG2DIM)) | (ly << INTERNAL_LOG2DIM) | lz;
  }
  bool has_child(size_t i) const { return ((child_mask[i / 64] >> (i % 64)) & 1ull) == 1ull; }
  T get(Coord c) const {  // :118-124
    const size_t i = child_index(c);
    return has_child(i) ? children[i]->get(c) : background;
  }
  void set(Coord c, T v) {  // :127-143 (new leaf filled with the tile value)
    const size_t i = child_index(c);
    if (!has_child(i)) {
      children[i].reset(new LeafNode<T>(aligned(c, LEAF_LOG2DIM), background));
      child_mask[i / 64] |= 1ull << (i % 64);
    }
    children[i]->set(c, v);
  }
  bool is_active(Coord c) const {  // :146-152
    const size_t i = child_index(c);
    return has_child(i) ? children[i]->is_active(c) : false;
  }
  void deactivate(Coord c) {  // :156-166
    const size_t i = child_index(c);
    if (has_child(i)) children[i]->set_inactive(c, background);
  }
  void remove_empty_leaves() {  // :231-247
    for (size_t i = 0; i < (size_t)INTERNAL_SIZE; ++i) {
      if (has_child(i) && children[i]->is_empty()) {
        children[i].reset();
        child_mask[i / 64] &= ~(1ull << (i % 64));
      }
    }
  }
  const LeafNode<T>* leaf_at(Coord c) const {
    const size_t i = child_index(c);
    return has_child(i) ? children[i].get() : nullptr;
  }
};

// ---------------------------------------------------------------- RootNode
// src/tree/root.rs:16-171
template <class T>
struct RootNode {
  std::unordered_map<Coord, InternalNode<T>, CoordHash> tiles;
  T background;

  explicit RootNode(T bg) : background(bg) {}

  T get(Coord c) const {  // :55-61
    auto it = tiles.find(aligned(c, INTERNAL_TOTAL_LOG2DIM));
    return it == tiles.end() ? background : it->second.get(c);
  }
  void set(Coord c, T v) {  // :64-72
    const Coord key = aligned(c, INTERNAL_TOTAL_LOG2DIM);
    auto it = tiles.find(key);
    if (it == tiles.end()) it = tiles.emplace(key, InternalNode<T>(key, background)).first;
    it->second.set(c, v);
  }
  bool is_active(Coord c) const {  // :75-80
    auto it = tiles.find(aligned(c, INTERNAL_TOTAL_LOG2DIM));
    return it != tiles.end() && it->second.is_active(c);
  }
  void deactivate(Coord c) {  // :142-147
    auto it = tiles.find(aligned(c, INTERNAL_TOTAL_LOG2DIM));
    if (it != tiles.end()) it->second.deactivate(c);
  }
  void remove_empty_leaves() {  // :161-165
    for (auto& kv : tiles) kv.second.remove_empty_leaves();
  }
  const LeafNode<T>* leaf_at(Coord c) const {
    auto it = tiles.find(aligned(c, INTERNAL_TOTAL_LOG2DIM));
    return it == tiles.end() ? nullptr : it->second.leaf_at(c);
  }
  template <class F>
  void for_each_leaf(F&& f) const {  // :83-85
    for (const auto& kv : tiles)
      for (size_t i = 0; i < (size_t)INTERNAL_SIZE; ++i)
        if (kv.second.has_child(i)) f(*kv.second.children[i]);
  }
  template <class F>
  void for_each_active(F&& f) const {  // :104-106
    for_each_leaf([&](const LeafNode<T>& l) { l.for_each_active(f); });
  }
  std::vector<Coord> leaf_origins() const {  // :93-95
    std::vector<Coord> o;
    for_each_leaf([&](const LeafNode<T>& l) { o.push_back(l.origin); });
    return o;
  }
  size_t leaf_count() const {
    size_t n = 0;
    for_each_leaf([&](const LeafNode<T>&) { ++n; });
    return n;
  }
  uint64_t active_voxel_count() const {
    uint64_t n = 0;
    for_each_leaf([&](const LeafNode<T>& l) { n += l.active_count(); });
    return n;
  }
};

// ---------------------------------------------------------------- transforms
// src/grid.rs:62-122 (legacy uniform transform)
struct VoxelTransform {
  double voxel_size = 1.0;
  double origin[3] = {0.0, 0.0, 0.0};
  static VoxelTransform uniform(double vs) {
    VoxelTransform t;
    t.voxel_size = vs;
    return t;
  }
  void world_to_index_f64(const double w[3], double out[3]) const {  // :114-121
    const double inv = 1.0 / voxel_size;
    out[0] = (w[0] - origin[0]) * inv;
    out[1] = (w[1] - origin[1]) * inv;
    out[2] = (w[2] - origin[2]) * inv;
  }
};

// src/grid.rs:131-296
struct AffineMap {
  double mat[4][4];
  double inv[4][4];
  double voxel_size;

  static AffineMap from_uniform_scale(double s) {  // :159-177
    AffineMap m;
    std::memset(&m, 0, sizeof(m));
    const double inv_s = 1.0 / s;
    m.mat[0][0] = m.mat[1][1] = m.mat[2][2] = s;
    m.mat[3][3] = 1.0;
    m.inv[0][0] = m.inv[1][1] = m.inv[2][2] = inv_s;
    m.inv[3][3] = 1.0;
    m.voxel_size = s;
    return m;
  }
  static AffineMap from_voxel_transform(const VoxelTransform& vt) {  // :180-199
    AffineMap m;
    std::memset(&m, 0, sizeof(m));
    const double s = vt.voxel_size, inv_s = 1.0 / s;
    for (int r = 0; r < 3; ++r) {
      m.mat[r][r] = s;
      m.mat[r][3] = vt.origin[r];
      m.inv[r][r] = inv_s;
      m.inv[r][3] = -vt.origin[r] * inv_s;  // -0.0 when the origin is +0.0
    }
    m.mat[3][3] = 1.0;
    m.inv[3][3] = 1.0;
    m.voxel_size = s;
    return m;
  }
  void index_to_world(Coord c, double out[3]) const {  // :248-257
    const double x = (double)c.x, y = (double)c.y, z = (double)c.z;
    for (int r = 0; r < 3; ++r) out[r] = mat[r][0] * x + mat[r][1] * y + mat[r][2] * z + mat[r][3];
  }
};

// src/grid.rs:48-55; numbering from src/nano/layout.rs:64-78.
enum GridClass : uint32_t { GC_UNKNOWN = 0, GC_LEVEL_SET = 1, GC_FOG_VOLUME = 2 };

// ---------------------------------------------------------------- Grid
// src/grid.rs:349-514
template <class T>
struct Grid {
  RootNode<T> tree;
  AffineMap affine;
  VoxelTransform transform;
  GridClass grid_class = GC_UNKNOWN;
  std::string name;

  Grid(T background, double voxel_size)  // :366-375
      : tree(background), affine(AffineMap::from_uniform_scale(voxel_size)),
        transform(VoxelTransform::uniform(voxel_size)) {}
  static Grid with_affine(T background, const AffineMap& m) {  // :378-388
    Grid g(background, 1.0);
    g.affine = m;
    g.transform = VoxelTransform::uniform(m.voxel_size);
    return g;
  }
  static Grid level_set(double voxel_size, double half_width) {  // :391-404
    Grid g((T)(half_width * voxel_size), voxel_size);
    g.name = "sdf";
    g.grid_class = GC_LEVEL_SET;
    return g;
  }
  void set_transform(const VoxelTransform& t) {  // :432-435
    affine = AffineMap::from_voxel_transform(t);
    transform = t;
  }
  T background() const { return tree.background; }
  T get(Coord c) const { return tree.get(c); }
  void set(Coord c, T v) { tree.set(c, v); }
  size_t leaf_count() const { return tree.leaf_count(); }
  uint64_t active_voxel_count() const { return tree.active_voxel_count(); }
};

// Deep copy of a grid with a value-type conversion (float -> double for the
// FP64 comparison).  Leaves, masks and inactive stored values are kept.
template <class D, class S>
Grid<D> convert_grid(const Grid<S>& g) {
  Grid<D> out((D)g.background(), 1.0);
  out.affine = g.affine;
  out.transform = g.transform;
  out.grid_class = g.grid_class;
  out.name = g.name;
  g.tree.for_each_leaf([&](const LeafNode<S>& l) {
    out.tree.set(l.origin, (D)0);
    auto* dst = const_cast<LeafNode<D>*>(out.tree.leaf_at(l.origin));
    for (int i = 0; i < LEAF_SIZE; ++i) dst->values[i] = (D)l.values[i];
    std::memcpy(dst->active, l.active, sizeof(dst->active));
  });
  return out;
}
template <class T>
Grid<T> clone_grid(const Grid<T>& g) { return convert_grid<T, T>(g); }

// ---------------------------------------------------------------- trilinear
// src/tools/interpolation.rs:11-48.  With T = float: f64 index, f32 weights
// and blend, no contraction.
template <class T>
T sample_trilinear(const Grid<T>& g, const double world[3]) {
  double idx[3];
  g.transform.world_to_index_f64(world, idx);
  const double ix = std::floor(idx[0]), iy = std::floor(idx[1]), iz = std::floor(idx[2]);
  const T fx = (T)(idx[0] - ix), fy = (T)(idx[1] - iy), fz = (T)(idx[2] - iz);
  const int32_t i = f64_to_i32(ix), j = f64_to_i32(iy), k = f64_to_i32(iz);
  const T c000 = g.get(Coord(i, j, k));
  const T c100 = g.get(Coord(i + 1, j, k));
  const T c010 = g.get(Coord(i, j + 1, k));
  const T c110 = g.get(Coord(i + 1, j + 1, k));
  const T c001 = g.get(Coord(i, j, k + 1));
  const T c101 = g.get(Coord(i + 1, j, k + 1));
  const T c011 = g.get(Coord(i, j + 1, k + 1));
  const T c111 = g.get(Coord(i + 1, j + 1, k + 1));
  const T one = (T)1.0;
  const T c00 = c000 * (one - fx) + c100 * fx;
  const T c01 = c001 * (one - fx) + c101 * fx;
  const T c10 = c010 * (one - fx) + c110 * fx;
  const T c11 = c011 * (one - fx) + c111 * fx;
  const T c0 = c00 * (one - fy) + c10 * fy;
  const T c1 = c01 * (one - fy) + c11 * fy;
  return c0 * (one - fz) + c1 * fz;
}

// ---------------------------------------------------------------- velocity
// src/tools/advection.rs:10-24
template <class T>
struct VelocityField {
  Grid<T> vx, vy, vz;
  void sample(const double pos[3], double out[3]) const {
    out[0] = (double)sample_trilinear(vx, pos);
    out[1] = (double)sample_trilinear(vy, pos);
    out[2] = (double)sample_trilinear(vz, pos);
  }
};

// src/tools/advection.rs:116-134: a uniform field carried in the background
// of three grids, each with two active voxels.
template <class T>
VelocityField<T> make_uniform_velocity(double vs, const float vel[3], Coord bmin, Coord bmax) {
  VelocityField<T> f{Grid<T>((T)vel[0], vs), Grid<T>((T)vel[1], vs), Grid<T>((T)vel[2], vs)};
  f.vx.set(bmin, (T)vel[0]);
  f.vy.set(bmin, (T)vel[1]);
  f.vz.set(bmin, (T)vel[2]);
  f.vx.set(bmax, (T)vel[0]);
  f.vy.set(bmax, (T)vel[1]);
  f.vz.set(bmax, (T)vel[2]);
  return f;
}

// ---------------------------------------------------------------- advection
// src/tools/advection.rs:32-56 (RK1 semi-Lagrangian, same topology).
template <class T, class Vel>
Grid<T> advect_level_set(const Grid<T>& grid, const Vel& velocity, double dt) {
  Grid<T> out = Grid<T>::with_affine(grid.background(), grid.affine);
  out.name = grid.name;
  out.grid_class = grid.grid_class;
  grid.tree.for_each_active([&](Coord c, T) {
    double world[3], v[3];
    grid.affine.index_to_world(c, world);
    velocity.sample(world, v);
    const double xp[3] = {world[0] - dt * v[0], world[1] - dt * v[1], world[2] - dt * v[2]};
    out.set(c, sample_trilinear(grid, xp));
  });
  return out;
}

// ---------------------------------------------------------------- band tools
// src/tools/level_set_track.rs:14-47
template <class T>
std::vector<Coord> find_zero_crossings(const Grid<T>& grid) {
  static const Coord OFF[6] = {Coord(1, 0, 0), Coord(-1, 0, 0), Coord(0, 1, 0),
                               Coord(0, -1, 0), Coord(0, 0, 1), Coord(0, 0, -1)};
  std::vector<Coord> out;
  grid.tree.for_each_active([&](Coord c, T val) {
    if (val == (T)0.0) {
      out.push_back(c);
      return;
    }
    for (const Coord& o : OFF) {
      const Coord nb = c + o;
      if (!grid.tree.is_active(nb)) continue;
      const T nv = grid.tree.get(nb);
      if (val * nv < (T)0.0) {
        out.push_back(c);
        break;
      }
    }
  });
  return out;
}

// src/tools/level_set_track.rs:53-68
template <class T>
void trim_narrow_band(Grid<T>& grid, float half_width) {
  const T vs = (T)grid.transform.voxel_size;
  const T threshold = (T)half_width * vs;
  std::vector<Coord> kill;
  grid.tree.for_each_active([&](Coord c, T v) {
    if (std::fabs(v) >= threshold) kill.push_back(c);
  });
  for (const Coord& c : kill) grid.tree.deactivate(c);
  grid.tree.remove_empty_leaves();
}

// src/tools/level_set_track.rs:80-114, literally: O(active x crossings).
template <class T>
void rebuild_narrow_band(Grid<T>& grid, float half_width) {
  const T vs = (T)grid.transform.voxel_size;
  const T band = (T)half_width * vs;
  const std::vector<Coord> crossings = find_zero_crossings(grid);
  if (crossings.empty()) return;  // :85-87
  std::vector<std::pair<Coord, T>> active;
  grid.tree.for_each_active([&](Coord c, T v) { active.emplace_back(c, v); });
  for (const auto& cv : active) {
    const Coord c = cv.first;
    const T sign = cv.second >= (T)0.0 ? (T)1.0 : (T)-1.0;
    T min_dist = std::numeric_limits<T>::max();
    for (const Coord& x : crossings) {
      const T dx = (T)(c.x - x.x), dy = (T)(c.y - x.y), dz = (T)(c.z - x.z);
      const T d = std::sqrt(dx * dx + dy * dy + dz * dz);
      min_dist = std::fmin(min_dist, d);
    }
    const T nv = sign * min_dist * vs;
    if (std::fabs(nv) < band)
      grid.tree.set(c, nv);
    else
      grid.tree.deactivate(c);
  }
  grid.tree.remove_empty_leaves();
}

// Builder-defined: the bounded-radius form of rebuild_narrow_band.  Only
// crossings with |dx|,|dy|,|dz| <= ceil(half_width) can give a distance below
// half_width (rounding of sqrt and of the product by vs is monotone), so the
// kept set and values equal the brute-force form exactly.  The minimum is
// taken over integer squared distances; sqrt is monotone and correctly
// rounded, so min(sqrt(d2)) == sqrt(min(d2)).
template <class T>
void rebuild_narrow_band_bounded(Grid<T>& grid, float half_width) {
  const T vs = (T)grid.transform.voxel_size;
  const T band = (T)half_width * vs;
  const std::vector<Coord> crossings = find_zero_crossings(grid);
  if (crossings.empty()) return;
  std::unordered_map<Coord, char, CoordHash> cset;
  cset.reserve(crossings.size() * 2);
  for (const Coord& x : crossings) cset.emplace(x, 1);
  const int R = (int)std::ceil((double)half_width);
  std::vector<std::pair<Coord, T>> active;
  grid.tree.for_each_active([&](Coord c, T v) { active.emplace_back(c, v); });
  for (const auto& cv : active) {
    const Coord c = cv.first;
    const T sign = cv.second >= (T)0.0 ? (T)1.0 : (T)-1.0;
    int64_t best = -1;
    for (int dx = -R; dx <= R; ++dx)
      for (int dy = -R; dy <= R; ++dy)
        for (int dz = -R; dz <= R; ++dz) {
          const int64_t d2 = (int64_t)dx * dx + (int64_t)dy * dy + (int64_t)dz * dz;
          if (best >= 0 && d2 >= best) continue;
          if (cset.count(Coord(c.x + dx, c.y + dy, c.z + dz))) best = d2;
        }
    bool keep = false;
    T nv = 0;
    if (best >= 0) {
      nv = sign * std::sqrt((T)best) * vs;
      keep = std::fabs(nv) < band;
    }
    if (keep)
      grid.tree.set(c, nv);
    else
      grid.tree.deactivate(c);
  }
  grid.tree.remove_empty_leaves();
}

// ---------------------------------------------------------------- marching cubes
// src/tools/volume_to_mesh.rs:11-40
struct Mesh {
  std::vector<std::array<double, 3>> vertices;
  std::vector<std::array<uint32_t, 3>> triangles;
  bool is_watertight() const {  // :28-39
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> cnt;
    for (const auto& t : triangles)
      for (int i = 0; i < 3; ++i) {
        uint32_t a = t[i], b = t[(i + 1) % 3];
        cnt[a < b ? std::make_pair(a, b) : std::make_pair(b, a)] += 1;
      }
    for (const auto& kv : cnt)
      if (kv.second != 2) return false;
    return true;
  }
};

struct CoordPairHash {
  size_t operator()(const std::pair<Coord, Coord>& p) const {
    return CoordHash()(p.first) * 31u + CoordHash()(p.second);
  }
};

// src/tools/volume_to_mesh.rs:104-201
template <class T>
void mc_process_cell(const Grid<T>& grid, Coord base, T iso, double vs, Mesh& mesh,
                     std::unordered_map<std::pair<Coord, Coord>, uint32_t, CoordPairHash>& cache) {
  Coord corners[8];
  T vals[8];
  for (int i = 0; i < 8; ++i) {
    corners[i] = base + Coord(MC_CORNER[i][0], MC_CORNER[i][1], MC_CORNER[i][2]);
    vals[i] = grid.get(corners[i]);
  }
  uint32_t cube = 0;
  for (int i = 0; i < 8; ++i)
    if (vals[i] < iso) cube |= 1u << i;
  const uint32_t flags = mc_edge_mask(cube);
  if (flags == 0) return;
  uint32_t ev[12] = {0};
  for (int e = 0; e < 12; ++e) {
    if (!(flags & (1u << e))) continue;
    const int a = MC_EDGE[e][0], b = MC_EDGE[e][1];
    const auto key = corners[a] < corners[b] ? std::make_pair(corners[a], corners[b])
                                             : std::make_pair(corners[b], corners[a]);
    auto it = cache.find(key);
    if (it != cache.end()) {
      ev[e] = it->second;
      continue;
    }
    double t;
    if (std::fabs(vals[b] - vals[a]) > (T)1e-10)
      t = (double)((iso - vals[a]) / (vals[b] - vals[a]));
    else
      t = 0.5;
    const double pa[3] = {(double)corners[a].x * vs, (double)corners[a].y * vs, (double)corners[a].z * vs};
    const double pb[3] = {(double)corners[b].x * vs, (double)corners[b].y * vs, (double)corners[b].z * vs};
    const std::array<double, 3> pos = {pa[0] + t * (pb[0] - pa[0]), pa[1] + t * (pb[1] - pa[1]),
                                       pa[2] + t * (pb[2] - pa[2])};
    const uint32_t idx = (uint32_t)mesh.vertices.size();
    mesh.vertices.push_back(pos);
    cache.emplace(key, idx);
    ev[e] = idx;
  }
  const char* tri = MC_TRI[cube];
  for (size_t t = 0; tri[t]; t += 3)
    mesh.triangles.push_back({ev[mc_hex_digit(tri[t])], ev[mc_hex_digit(tri[t + 1])], ev[mc_hex_digit(tri[t + 2])]});
}

// The processed tiles: leaf origins and their 6 face neighbours, sorted and
// de-duplicated (src/tools/volume_to_mesh.rs:57-74).
template <class T>
std::vector<Coord> mc_tiles(const Grid<T>& grid) {
  std::vector<Coord> o;
  for (const Coord& c : grid.tree.leaf_origins()) {
    o.push_back(c);
    const Coord d[6] = {Coord(-8, 0, 0), Coord(8, 0, 0), Coord(0, -8, 0), Coord(0, 8, 0), Coord(0, 0, -8), Coord(0, 0, 8)};
    for (const Coord& dd : d) o.push_back(c + dd);
  }
  std::sort(o.begin(), o.end());
  o.erase(std::unique(o.begin(), o.end()), o.end());
  return o;
}

// src/tools/volume_to_mesh.rs:46-101
template <class T>
Mesh volume_to_mesh(const Grid<T>& grid, T iso) {
  Mesh mesh;
  std::unordered_map<std::pair<Coord, Coord>, uint32_t, CoordPairHash> cache;
  const double vs = grid.transform.voxel_size;
  for (const Coord& o : mc_tiles(grid))
    for (int lx = 0; lx < 8; ++lx)
      for (int ly = 0; ly < 8; ++ly)
        for (int lz = 0; lz < 8; ++lz) mc_process_cell(grid, o + Coord(lx, ly, lz), iso, vs, mesh, cache);
  return mesh;
}

// ---------------------------------------------------------------- NanoGrid
// src/nano/layout.rs:13-52: a 320-byte header then 2128-byte leaves sorted
// by origin.  Offsets are the repr(C) ones.
constexpr size_t NANO_HEADER_SIZE = 320;
constexpr size_t NANO_LEAF_SIZE = 2128;
constexpr uint32_t NANO_VERSION = 1;  // src/nano/layout.rs:10

struct NanoHeaderView {
  char magic[4];
  uint32_t version, grid_class;
  double voxel_size;
  float background;
  uint32_t leaf_count;
  int32_t bbox_min[3], bbox_max[3];
  double affine_mat[4][4], affine_inv[4][4];
};

inline void nano_read_header(const uint8_t* b, NanoHeaderView& h) {
  std::memcpy(h.magic, b + 0, 4);
  std::memcpy(&h.version, b + 4, 4);
  std::memcpy(&h.grid_class, b + 8, 4);
  std::memcpy(&h.voxel_size, b + 16, 8);
  std::memcpy(&h.background, b + 24, 4);
  std::memcpy(&h.leaf_count, b + 28, 4);
  std::memcpy(h.bbox_min, b + 32, 12);
  std::memcpy(h.bbox_max, b + 44, 12);
  std::memcpy(h.affine_mat, b + 56, 128);
  std::memcpy(h.affine_inv, b + 184, 128);
}

// src/nano/layout.rs:89-170
inline std::vector<uint8_t> nano_from_grid(const Grid<float>& grid) {
  struct L {
    Coord o;
    const LeafNode<float>* p;
  };
  std::vector<L> leaves;
  grid.tree.for_each_leaf([&](const LeafNode<float>& l) { leaves.push_back({l.origin, &l}); });
  std::sort(leaves.begin(), leaves.end(), [](const L& a, const L& b) { return a.o < b.o; });
  int32_t bmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  grid.tree.for_each_active([&](Coord c, float) {
    bmin[0] = std::min(bmin[0], c.x); bmin[1] = std::min(bmin[1], c.y); bmin[2] = std::min(bmin[2], c.z);
    bmax[0] = std::max(bmax[0], c.x); bmax[1] = std::max(bmax[1], c.y); bmax[2] = std::max(bmax[2], c.z);
  });
  if (bmin[0] == INT32_MAX) {
    for (int i = 0; i < 3; ++i) bmin[i] = bmax[i] = 0;
  }
  std::vector<uint8_t> out(NANO_HEADER_SIZE + NANO_LEAF_SIZE * leaves.size(), 0);
  uint8_t* b = out.data();
  std::memcpy(b, "NANO", 4);
  const uint32_t version = NANO_VERSION, gc = (uint32_t)grid.grid_class, lc = (uint32_t)leaves.size();
  const double vs = grid.affine.voxel_size;
  const float bg = grid.background();
  std::memcpy(b + 4, &version, 4);
  std::memcpy(b + 8, &gc, 4);
  std::memcpy(b + 16, &vs, 8);
  std::memcpy(b + 24, &bg, 4);
  std::memcpy(b + 28, &lc, 4);
  std::memcpy(b + 32, bmin, 12);
  std::memcpy(b + 44, bmax, 12);
  std::memcpy(b + 56, grid.affine.mat, 128);
  std::memcpy(b + 184, grid.affine.inv, 128);
  for (size_t i = 0; i < leaves.size(); ++i) {
    uint8_t* lp = b + NANO_HEADER_SIZE + i * NANO_LEAF_SIZE;
    const int32_t o[3] = {leaves[i].o.x, leaves[i].o.y, leaves[i].o.z};
    std::memcpy(lp, o, 12);
    std::memcpy(lp + 16, leaves[i].p->active, 64);
    std::memcpy(lp + 80, leaves[i].p->values.data(), 2048);
  }
  return out;
}

// src/nano/layout.rs:202-251; returns "" when valid, else the error text.
inline std::string nano_validate(const uint8_t* b, size_t len) {
  if (len < NANO_HEADER_SIZE) return "Buffer too small for NanoHeader";
  NanoHeaderView h;
  nano_read_header(b, h);
  if (std::memcmp(h.magic, "NANO", 4) != 0) return "Invalid magic";
  if (h.version != NANO_VERSION) return "Unsupported version";
  const size_t expected = NANO_HEADER_SIZE + NANO_LEAF_SIZE * (size_t)h.leaf_count;
  if (len < expected) return "Buffer too small";
  for (uint32_t i = 1; i < h.leaf_count; ++i) {
    int32_t p[3], c[3];
    std::memcpy(p, b + NANO_HEADER_SIZE + (i - 1) * NANO_LEAF_SIZE, 12);
    std::memcpy(c, b + NANO_HEADER_SIZE + i * NANO_LEAF_SIZE, 12);
    if (!(Coord(p[0], p[1], p[2]) < Coord(c[0], c[1], c[2]))) return "Leaves not sorted";
  }
  return "";
}

// src/nano/layout.rs:254-285 (active voxels only are re-inserted)
inline Grid<float> nano_to_grid(const uint8_t* b) {
  NanoHeaderView h;
  nano_read_header(b, h);
  AffineMap m;
  std::memcpy(m.mat, h.affine_mat, 128);
  std::memcpy(m.inv, h.affine_inv, 128);
  m.voxel_size = h.voxel_size;
  Grid<float> g = Grid<float>::with_affine(h.background, m);
  g.grid_class = (h.grid_class == 1) ? GC_LEVEL_SET : (h.grid_class == 2 ? GC_FOG_VOLUME : GC_UNKNOWN);
  for (uint32_t i = 0; i < h.leaf_count; ++i) {
    const uint8_t* lp = b + NANO_HEADER_SIZE + (size_t)i * NANO_LEAF_SIZE;
    int32_t o[3];
    uint64_t mask[8];
    std::memcpy(o, lp, 12);
    std::memcpy(mask, lp + 16, 64);
    for (int w = 0; w < 8; ++w) {
      uint64_t word = mask[w];
      while (word) {
        const int bit = __builtin_ctzll(word);
        word &= word - 1;
        const int off = w * 64 + bit;
        float v;
        std::memcpy(&v, lp + 80 + 4 * off, 4);
        g.set(Coord(o[0] + (off >> 6), o[1] + ((off >> 3) & 7), o[2] + (off & 7)), v);
      }
    }
  }
  return g;
}

// src/nano/lookup.rs:58-79 (binary search) and :11-31 (get / is_active)
inline const uint8_t* nano_find_leaf(const uint8_t* b, Coord c) {
  NanoHeaderView h;
  nano_read_header(b, h);
  const Coord t = aligned(c, LEAF_LOG2DIM);
  size_t lo = 0, hi = h.leaf_count;
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    const uint8_t* lp = b + NANO_HEADER_SIZE + mid * NANO_LEAF_SIZE;
    int32_t o[3];
    std::memcpy(o, lp, 12);
    const Coord mo(o[0], o[1], o[2]);
    if (mo < t)
      lo = mid + 1;
    else if (mo == t)
      return lp;
    else
      hi = mid;
  }
  return nullptr;
}
inline float nano_get(const uint8_t* b, Coord c) {
  const uint8_t* lp = nano_find_leaf(b, c);
  float v;
  if (!lp) {
    std::memcpy(&v, b + 24, 4);
    return v;
  }
  std::memcpy(&v, lp + 80 + 4 * offset_in_tile(c, LEAF_LOG2DIM), 4);
  return v;
}
inline bool nano_is_active(const uint8_t* b, Coord c) {
  const uint8_t* lp = nano_find_leaf(b, c);
  if (!lp) return false;
  const size_t off = offset_in_tile(c, LEAF_LOG2DIM);
  uint64_t w;
  std::memcpy(&w, lp + 16 + 8 * (off / 64), 8);
  return ((w >> (off % 64)) & 1ull) == 1ull;
}

// ---------------------------------------------------------------- fixtures
// src/tools/sphere.rs:17-59
template <class T>
Grid<T> make_level_set_sphere(double radius, const double center[3], double voxel_size, double half_width) {
  const double band = half_width * voxel_size;
  const T bg = (T)band;
  Grid<T> grid(bg, voxel_size);
  grid.grid_class = GC_LEVEL_SET;
  grid.name = "sdf";
  grid.set_transform(VoxelTransform::uniform(voxel_size));
  const double extent = radius + band;
  const double inv = 1.0 / voxel_size;
  const int32_t imin = f64_to_i32(std::floor((center[0] - extent) * inv));
  const int32_t jmin = f64_to_i32(std::floor((center[1] - extent) * inv));
  const int32_t kmin = f64_to_i32(std::floor((center[2] - extent) * inv));
  const int32_t imax = f64_to_i32(std::ceil((center[0] + extent) * inv));
  const int32_t jmax = f64_to_i32(std::ceil((center[1] + extent) * inv));
  const int32_t kmax = f64_to_i32(std::ceil((center[2] + extent) * inv));
  for (int32_t i = imin; i <= imax; ++i) {
    const double wx = (double)i * voxel_size - center[0];
    for (int32_t j = jmin; j <= jmax; ++j) {
      const double wy = (double)j * voxel_size - center[1];
      for (int32_t k = kmin; k <= kmax; ++k) {
        const double wz = (double)k * voxel_size - center[2];
        const double dist = std::sqrt(wx * wx + wy * wy + wz * wz);
        const T sdf = (T)(dist - radius);
        if (std::fabs(sdf) < bg) grid.set(Coord(i, j, k), sdf);
      }
    }
  }
  return grid;
}

// src/tools/csg.rs:32-78.  `combine` is min (union), max (intersection) or
// max(a, -b) (difference).  Rust's f32::min/max return the other operand for
// a NaN; for equal operands (+0 vs -0) this restatement returns the first.
enum CsgOp { CSG_UNION = 0, CSG_INTERSECTION = 1, CSG_DIFFERENCE = 2 };
inline float csg_min(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return b < a ? b : a;
}
inline float csg_max(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return b > a ? b : a;
}
inline float csg_apply(CsgOp op, float a, float b) {
  switch (op) {
    case CSG_UNION: return csg_min(a, b);
    case CSG_INTERSECTION: return csg_max(a, b);
    default: return csg_max(a, -b);
  }
}
inline Grid<float> csg_combine(const Grid<float>& a, const Grid<float>& b, CsgOp op) {
  const float bg_out = csg_apply(op, a.background(), b.background());
  const double vs = a.transform.voxel_size;
  Grid<float> out(bg_out, vs);
  out.grid_class = GC_LEVEL_SET;
  out.set_transform(a.transform);
  std::vector<Coord> all = a.tree.leaf_origins();
  const std::vector<Coord> ob = b.tree.leaf_origins();
  all.insert(all.end(), ob.begin(), ob.end());
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  for (const Coord& o : all)
    for (int lx = 0; lx < 8; ++lx)
      for (int ly = 0; ly < 8; ++ly)
        for (int lz = 0; lz < 8; ++lz) {
          const Coord c = o + Coord(lx, ly, lz);
          const float va = a.get(c), vb = b.get(c);
          const float comb = csg_apply(op, va, vb);
          const bool either = a.tree.is_active(c) || b.tree.is_active(c);
          if (either || std::fabs(comb - bg_out) > 1e-12f) out.set(c, comb);
        }
  return out;
}

}  // namespace vo
