// Refinement tree and active-leaf snapshots; see topology.hpp for the
// reference citations.
#include "topology.hpp"

#include <algorithm>
#include <string>

namespace gh {

Grid::Grid(int dim, const int64_t* root_counts) : D(dim) {
  if (dim < 1 || dim > kMaxDim) throw std::invalid_argument("dimension must be 1, 2 or 3");
  int64_t total = 1;
  for (int a = 0; a < D; ++a) {
    if (root_counts[a] < 1) throw std::invalid_argument("root_counts must be positive");
    root[a] = root_counts[a];
    total *= root_counts[a];
  }
  // Root cells are stored in mixed-radix order with axis 1 fastest.
  for (int64_t id = 0; id < total; ++id) {
    std::array<int32_t, kMaxDim> lv{{0, 0, 0}};
    std::array<int64_t, kMaxDim> c{{0, 0, 0}};
    int64_t rem = id;
    for (int a = 0; a < D; ++a) {
      c[a] = rem % root[a];
      rem /= root[a];
    }
    level.push_back(lv);
    coord.push_back(c);
    parent.push_back(-1);
    first_child.push_back(-1);
    split.push_back(-1);
    active.push_back(1);
  }
}

int64_t Grid::refine(int64_t cell, int axis) {
  if (cell < 0 || cell >= cell_count()) throw std::invalid_argument("cell out of range");
  if (axis < 0 || axis >= D) throw std::invalid_argument("axis out of range");
  if (!active[cell]) throw std::invalid_argument("only active leaves can be refined");
  if (level[cell][axis] >= 40) throw std::invalid_argument("refinement level limit reached");
  int64_t first = cell_count();
  for (int off = 0; off < 2; ++off) {
    auto lv = level[cell];
    auto c = coord[cell];
    lv[axis] += 1;
    c[axis] = 2 * c[axis] + off;
    level.push_back(lv);
    coord.push_back(c);
    parent.push_back(cell);
    first_child.push_back(-1);
    split.push_back(-1);
    active.push_back(1);
  }
  first_child[cell] = first;
  split[cell] = int8_t(axis);
  active[cell] = 0;
  return first;
}

std::vector<int64_t> Grid::active_leaves() const {
  std::vector<int64_t> out;
  for (int64_t c = 0; c < cell_count(); ++c)
    if (active[c]) out.push_back(c);
  return out;
}

namespace {

using u128 = unsigned __int128;

struct MortonKey {
  u128 code;
  std::array<int32_t, kMaxDim> level;
  int64_t cell;
};

}  // namespace

Snapshot::Snapshot(const Grid& g, const std::vector<int64_t>& active) : grid(&g), D(g.D) {
  if (active.empty()) throw std::invalid_argument("grid snapshot must contain at least one active leaf");
  for (int64_t leaf : active) {
    if (leaf < 0 || leaf >= g.cell_count()) throw std::invalid_argument("leaf id out of range");
    for (int a = 0; a < D; ++a) max_level[a] = std::max(max_level[a], g.level[leaf][a]);
  }
  // Bits needed on the common finest lattice (maximum over axes).
  int bit_count = 0;
  for (int a = 0; a < D; ++a) {
    int64_t extent = g.root[a] << max_level[a];
    int64_t value = extent - 1;
    int bits = 0;
    while (value > 0) {
      ++bits;
      value >>= 1;
    }
    bit_count = std::max(bit_count, bits);
  }
  if (bit_count * D > 126) throw std::invalid_argument("Morton key exceeds 126 bits");
  std::vector<MortonKey> keys;
  keys.reserve(active.size());
  for (int64_t leaf : active) {
    MortonKey k;
    k.code = 0;
    for (int bit = bit_count - 1; bit >= 0; --bit)
      for (int a = D - 1; a >= 0; --a) {
        uint64_t lower = uint64_t(g.coord[leaf][a]) << (max_level[a] - g.level[leaf][a]);
        k.code = (k.code << 1) | u128((lower >> bit) & 1u);
      }
    k.level = g.level[leaf];
    k.cell = leaf;
    keys.push_back(k);
  }
  std::sort(keys.begin(), keys.end(), [&](const MortonKey& l, const MortonKey& r) {
    if (l.code != r.code) return l.code < r.code;
    for (int a = 0; a < D; ++a)
      if (l.level[a] != r.level[a]) return l.level[a] < r.level[a];
    return l.cell < r.cell;
  });
  leaves.reserve(keys.size());
  leaf_index.reserve(keys.size() * 2);
  for (size_t i = 0; i < keys.size(); ++i) {
    if (leaf_index.count(keys[i].cell)) throw std::invalid_argument("snapshot active leaves must be unique");
    leaves.push_back(keys[i].cell);
    leaf_index[keys[i].cell] = int64_t(i);
  }
  // Tree cells: ancestors-or-self of every active leaf.
  tree.reserve(leaves.size() * 3);
  for (int64_t leaf : leaves) {
    int64_t cur = leaf;
    while (cur != -1) {
      CellKey key{g.level[cur], g.coord[cur]};
      auto it = tree.find(key);
      if (it != tree.end()) {
        if (it->second != cur) throw std::invalid_argument("snapshot tree cells must have unique logical signatures");
        break;
      }
      tree.emplace(key, cur);
      cur = g.parent[cur];
    }
  }
  // Boundary faces: leaf, axis, side (lower then upper) without a covering neighbour.
  for (size_t i = 0; i < leaves.size(); ++i)
    for (int a = 0; a < D; ++a)
      for (int side = 0; side < 2; ++side)
        if (covering_neighbor(leaves[i], a, side) == -1) {
          boundary_leaf.push_back(int64_t(i));
          boundary_axis.push_back(a);
          boundary_side.push_back(side);
        }
  // Interfaces: leaf, axis, opposite active leaves on the upper side.
  for (size_t i = 0; i < leaves.size(); ++i)
    for (int a = 0; a < D; ++a)
      for (int64_t other : opposite_leaves(leaves[i], a, 1)) {
        iface_minus.push_back(int64_t(i));
        iface_plus.push_back(leaf_index.at(other));
        iface_axis.push_back(a);
      }
}

bool Snapshot::in_tree(int64_t cell) const {
  auto it = tree.find(CellKey{grid->level[cell], grid->coord[cell]});
  return it != tree.end() && it->second == cell;
}

int64_t Snapshot::covering_neighbor(int64_t cell, int axis, int side) const {
  const Grid& g = *grid;
  int64_t cur = cell;
  while (true) {
    CellKey key{g.level[cur], g.coord[cur]};
    int64_t extent = g.root[axis] << g.level[cur][axis];
    int64_t nc = key.coord[axis] + (side == 0 ? -1 : 1);
    if (nc >= 0 && nc < extent) {
      key.coord[axis] = nc;
      auto it = tree.find(key);
      if (it != tree.end()) return it->second;
    }
    int64_t par = g.parent[cur];
    if (par == -1) return -1;
    // A child touches the queried side of its parent unless the parent was
    // split normal to that side and the child is the other half.
    if (g.split[par] == axis && (g.coord[cur][axis] & 1) != side) return -1;
    cur = par;
  }
}

bool Snapshot::tangential_overlap(int64_t a, int64_t b, int face_axis) const {
  const Grid& g = *grid;
  for (int ax = 0; ax < D; ++ax) {
    if (ax == face_axis) continue;
    int32_t lvl = std::max(g.level[a][ax], g.level[b][ax]);
    int64_t al = g.coord[a][ax] << (lvl - g.level[a][ax]);
    int64_t au = (g.coord[a][ax] + 1) << (lvl - g.level[a][ax]);
    int64_t bl = g.coord[b][ax] << (lvl - g.level[b][ax]);
    int64_t bu = (g.coord[b][ax] + 1) << (lvl - g.level[b][ax]);
    if (al >= bu || bl >= au) return false;
  }
  return true;
}

void Snapshot::collect_opposite(int64_t source, int axis, int side, int64_t candidate,
                                std::vector<int64_t>& out) const {
  const Grid& g = *grid;
  if (!tangential_overlap(source, candidate, axis)) return;
  if (is_leaf(candidate)) {
    out.push_back(candidate);
    return;
  }
  int sp = g.split[candidate];
  int64_t first = g.first_child[candidate];
  if (sp < 0 || first < 0) return;
  if (sp == axis) {
    int64_t child = first + (side == 0 ? 1 : 0);
    if (!in_tree(child)) throw std::invalid_argument("snapshot tree is missing an active face child");
    collect_opposite(source, axis, side, child, out);
    return;
  }
  for (int64_t child = first; child < first + 2; ++child) {
    if (!in_tree(child)) continue;
    collect_opposite(source, axis, side, child, out);
  }
}

std::vector<int64_t> Snapshot::opposite_leaves(int64_t cell, int axis, int side) const {
  std::vector<int64_t> out;
  int64_t candidate = covering_neighbor(cell, axis, side);
  if (candidate == -1) return out;
  collect_opposite(cell, axis, side, candidate, out);
  if (out.size() <= 1) return out;
  const Grid& g = *grid;
  std::array<int32_t, kMaxDim> cmp{{0, 0, 0}};
  for (int a = 0; a < D; ++a) {
    cmp[a] = g.level[cell][a];
    for (int64_t l : out) cmp[a] = std::max(cmp[a], g.level[l][a]);
  }
  auto key = [&](int64_t l) {
    std::array<int64_t, kMaxDim> k{{0, 0, 0}};
    for (int a = 0; a < D; ++a) k[a] = (a == axis) ? 0 : (g.coord[l][a] << (cmp[a] - g.level[l][a]));
    return k;
  };
  std::stable_sort(out.begin(), out.end(), [&](int64_t x, int64_t y) { return key(x) < key(y); });
  return out;
}

}  // namespace gh
