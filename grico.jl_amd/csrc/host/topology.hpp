// Binary (one-axis) dyadic refinement tree on a Cartesian root grid, and the
// frozen "snapshot" of an active-leaf set: Morton order, boundary faces and
// interface pairs.
//
// Follows src/topology.jl: CartesianGrid (96-130), refine! / midpoint split
// (src/refinement.jl:47-80), Morton snapshot order (250-300: compare lower
// corners scaled to the finest lattice bit by bit from the top, axis D first,
// ties by level tuple then cell id), boundary enumeration (1502-1521: leaf,
// axis, side), interface enumeration (1523-1545: leaf, axis, opposite leaves on
// the upper side) and the covering-neighbour / descend search (1596-1663).
#pragma once
#include <array>
#include <cstdint>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace gh {

constexpr int kMaxDim = 3;

struct Grid {
  int D = 0;
  std::array<int64_t, kMaxDim> root{{1, 1, 1}};
  std::vector<std::array<int32_t, kMaxDim>> level;
  std::vector<std::array<int64_t, kMaxDim>> coord;
  std::vector<int64_t> parent;
  std::vector<int64_t> first_child;
  std::vector<int8_t> split;  // -1: not expanded, else split axis (0-based)
  std::vector<uint8_t> active;

  Grid(int dim, const int64_t* root_counts);
  int64_t cell_count() const { return int64_t(parent.size()); }
  // Split an active leaf at its midpoint normal to `axis`; returns the first
  // child id (the second child is first + 1).
  int64_t refine(int64_t cell, int axis);
  std::vector<int64_t> active_leaves() const;
};

struct CellKey {
  std::array<int32_t, kMaxDim> level;
  std::array<int64_t, kMaxDim> coord;
  bool operator==(const CellKey& o) const { return level == o.level && coord == o.coord; }
};

struct CellKeyHash {
  size_t operator()(const CellKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int a = 0; a < kMaxDim; ++a) {
      h = (h ^ uint64_t(k.level[a])) * 1099511628211ull;
      h = (h ^ uint64_t(k.coord[a])) * 1099511628211ull;
    }
    return size_t(h ^ (h >> 29));
  }
};

// A frozen active-leaf set of a grid (the reference's GridSnapshot).
struct Snapshot {
  const Grid* grid = nullptr;
  int D = 0;
  std::vector<int64_t> leaves;        // Morton order
  std::unordered_map<int64_t, int64_t> leaf_index;  // cell id -> index in `leaves`
  std::array<int32_t, kMaxDim> max_level{{0, 0, 0}};
  // Tree cells (ancestors-or-self of active leaves) keyed by logical signature.
  std::unordered_map<CellKey, int64_t, CellKeyHash> tree;
  std::vector<int64_t> boundary_leaf;  // leaf index
  std::vector<int32_t> boundary_axis;  // 0-based
  std::vector<int32_t> boundary_side;  // 0 lower, 1 upper
  std::vector<int64_t> iface_minus;    // leaf index
  std::vector<int64_t> iface_plus;     // leaf index
  std::vector<int32_t> iface_axis;     // 0-based

  Snapshot(const Grid& g, const std::vector<int64_t>& active);

  bool is_leaf(int64_t cell) const { return leaf_index.count(cell) != 0; }
  bool in_tree(int64_t cell) const;
  int64_t covering_neighbor(int64_t cell, int axis, int side) const;
  void collect_opposite(int64_t source, int axis, int side, int64_t candidate,
                        std::vector<int64_t>& out) const;
  std::vector<int64_t> opposite_leaves(int64_t cell, int axis, int side) const;
  bool tangential_overlap(int64_t a, int64_t b, int face_axis) const;
};

}  // namespace gh
