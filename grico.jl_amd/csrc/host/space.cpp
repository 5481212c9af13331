// Continuity compiler: see space.hpp for the reference citations.
#include "space.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <string>
#include <tuple>
#include <unordered_map>

#include "polys.hpp"

namespace gh {

namespace {

int64_t box_count(int D, const int32_t* deg) {
  int64_t n = 1;
  for (int a = 0; a < D; ++a) n *= int64_t(deg[a]) + 1;
  return n;
}

bool trunk_active(int D, const int32_t* deg, const int32_t* mode) {
  int32_t limit = 0, order = 0;
  for (int a = 0; a < D; ++a) {
    limit = std::max(limit, deg[a]);
    order += mode[a] <= 1 ? 0 : mode[a];
  }
  return order <= limit;
}

// Cached local pattern of one degree tuple.
struct Pattern {
  std::array<int32_t, kMaxDim> deg{{0, 0, 0}};
  std::array<int32_t, kMaxDim> shape{{1, 1, 1}};
  std::vector<std::array<int32_t, kMaxDim>> modes;  // local order
  std::vector<int32_t> mode_box;                    // box index per local mode
  std::vector<int32_t> box_coupled_rank;            // per box index: variable rank or -1
  int32_t coupled_count = 0;
  // face_modes[axis][side]: box indices of coupled boundary modes with
  // mode[axis] == side, in local order.
  std::array<std::array<std::vector<int32_t>, 2>, kMaxDim> face_modes;
};

int32_t flatten(const std::array<int32_t, kMaxDim>& mode, const std::array<int32_t, kMaxDim>& shape,
                int D) {
  int32_t idx = 0, stride = 1;
  for (int a = 0; a < D; ++a) {
    idx += mode[a] * stride;
    stride *= shape[a];
  }
  return idx;
}

std::array<int32_t, kMaxDim> unflatten(int32_t idx, const std::array<int32_t, kMaxDim>& shape, int D) {
  std::array<int32_t, kMaxDim> m{{0, 0, 0}};
  for (int a = 0; a < D; ++a) {
    m[a] = idx % shape[a];
    idx /= shape[a];
  }
  return m;
}

bool requires_coupling(int D, int cg_mask, const std::array<int32_t, kMaxDim>& mode) {
  for (int a = 0; a < D; ++a)
    if ((cg_mask >> a & 1) && mode[a] <= 1) return true;
  return false;
}

Pattern make_pattern(int basis, int D, int cg_mask, const std::array<int32_t, kMaxDim>& deg) {
  Pattern p;
  p.deg = deg;
  for (int a = 0; a < D; ++a) p.shape[a] = deg[a] + 1;
  p.modes = basis_modes(basis, D, deg.data());
  int64_t nbox = box_count(D, deg.data());
  p.box_coupled_rank.assign(size_t(nbox), -1);
  for (const auto& m : p.modes) {
    int32_t b = flatten(m, p.shape, D);
    p.mode_box.push_back(b);
    bool boundary = false;
    for (int a = 0; a < D; ++a) boundary = boundary || m[a] <= 1;
    if (!boundary || !requires_coupling(D, cg_mask, m)) continue;
    p.box_coupled_rank[b] = p.coupled_count++;
    for (int a = 0; a < D; ++a)
      if ((cg_mask >> a & 1) && m[a] <= 1) p.face_modes[a][m[a]].push_back(b);
  }
  return p;
}

// Sparse row over variables or dofs, kept in insertion order like the
// reference's _ConstraintRow.
//
// Cancellation test: the reference drops coefficients below 4 eps times the
// row's current largest coefficient (src/continuity.jl:687-698, 1213-1223).
// After substitutions that cancel a row completely, that scale is the noise
// itself, so a residue of a few ulps of the substituted terms survives and
// becomes a spurious pivot that over-constrains the space (two-level hanging
// faces of p = 2 show it: one coarse edge bubble is forced to zero). `mag`
// keeps the largest magnitude ever added to the row, and the tolerance uses it
// as well; rows that never cancelled are cleaned exactly as in the reference.
struct Row {
  std::vector<int64_t> var;
  std::vector<double> coef;
  double mag = 0.0;
  void add(int64_t v, double c) {
    mag = std::max(mag, std::fabs(c));
    for (size_t i = 0; i < var.size(); ++i)
      if (var[i] == v) {
        coef[i] += c;
        return;
      }
    var.push_back(v);
    coef.push_back(c);
  }
  void remove_at(size_t i) {
    var.erase(var.begin() + long(i));
    coef.erase(coef.begin() + long(i));
  }
  bool empty() const { return var.empty(); }
  void cleanup() {
    double scale = 0.0;
    for (double c : coef) scale = std::max(scale, std::fabs(c));
    double tol = 4.0 * std::numeric_limits<double>::epsilon() * std::max(std::max(scale, mag), 1.0);
    size_t i = 0;
    while (i < var.size()) {
      if (std::fabs(coef[i]) <= tol)
        remove_at(i);
      else
        ++i;
    }
  }
};

struct UnionFind {
  std::vector<int64_t> parent;
  explicit UnionFind(int64_t n) : parent(size_t(n)) {
    for (int64_t i = 0; i < n; ++i) parent[size_t(i)] = i;
  }
  int64_t root(int64_t v) {
    int64_t r = v;
    while (parent[size_t(r)] != r) r = parent[size_t(r)];
    while (parent[size_t(v)] != v) {
      int64_t nx = parent[size_t(v)];
      parent[size_t(v)] = r;
      v = nx;
    }
    return r;
  }
  // Keep the smaller id as representative.
  void unite(int64_t a, int64_t b) {
    int64_t ra = root(a), rb = root(b);
    if (ra == rb) return;
    if (ra < rb)
      parent[size_t(rb)] = ra;
    else
      parent[size_t(ra)] = rb;
  }
};

void interval_level_and_coord(int64_t lower, int64_t upper, int finest, int& level, int64_t& coord) {
  int64_t len = upper - lower;
  if (lower < 0 || len <= 0 || (len & (len - 1)) != 0)
    throw std::invalid_argument("dyadic interval length must be a power of two");
  int shift = __builtin_ctzll(uint64_t(len));
  if (shift > finest) throw std::invalid_argument("dyadic interval is coarser than the finest lattice");
  if (lower % len != 0) throw std::invalid_argument("dyadic interval lower bound is misaligned");
  level = finest - shift;
  coord = lower >> shift;
}

}  // namespace

std::vector<std::array<int32_t, kMaxDim>> basis_modes(int basis, int D, const int32_t* degrees) {
  std::vector<std::array<int32_t, kMaxDim>> out;
  std::array<int32_t, kMaxDim> shape{{1, 1, 1}};
  for (int a = 0; a < D; ++a) {
    if (degrees[a] < 0) throw std::invalid_argument("degrees must be non-negative");
    shape[a] = degrees[a] + 1;
  }
  int64_t n = box_count(D, degrees);
  for (int64_t i = 0; i < n; ++i) {
    auto m = unflatten(int32_t(i), shape, D);
    if (basis == kTrunk && !trunk_active(D, degrees, m.data())) continue;
    out.push_back(m);
  }
  return out;
}

std::vector<double> affine_restriction_matrix(int source_degree, int target_degree, int delta,
                                              int64_t relative) {
  if (target_degree < source_degree)
    throw std::invalid_argument("target degree must represent the source degree exactly");
  double scale = std::ldexp(1.0, -delta);
  double shift = delta == 0 ? 0.0 : std::fma(scale, 2.0 * double(relative) + 1.0, -1.0);
  int n = target_degree + 1, ns = source_degree + 1;
  const size_t nn = size_t(n), nns = size_t(ns);
  std::vector<double> pts(nn), wts(nn), vals(nn);
  gauss_legendre(n, pts.data(), wts.data());
  // A (n x n): target basis at target points; S (n x ns): source basis at the
  // pulled-back points. Solve A X = S by LU with partial pivoting.
  std::vector<double> A(nn * nn), S(nn * nns);
  std::vector<double> sv(nns);
  for (int i = 0; i < n; ++i) {
    fe_basis(pts[size_t(i)], target_degree, vals.data(), nullptr);
    for (int j = 0; j < n; ++j) A[size_t(i * n + j)] = vals[size_t(j)];
    double xi = std::fma(scale, pts[size_t(i)], shift);
    fe_basis(xi, source_degree, sv.data(), nullptr);
    for (int j = 0; j < ns; ++j) S[size_t(i * ns + j)] = sv[size_t(j)];
  }
    for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = std::fabs(A[size_t(k * n + k)]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[size_t(i * n + k)]) > best) {
        best = std::fabs(A[size_t(i * n + k)]);
        piv = i;
      }
    if (best == 0.0) throw std::runtime_error("singular restriction collocation matrix");
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[size_t(k * n + j)], A[size_t(piv * n + j)]);
      for (int j = 0; j < ns; ++j) std::swap(S[size_t(k * ns + j)], S[size_t(piv * ns + j)]);
    }
    for (int i = k + 1; i < n; ++i) {
      double f = A[size_t(i * n + k)] / A[size_t(k * n + k)];
      A[size_t(i * n + k)] = f;
      for (int j = k + 1; j < n; ++j) A[size_t(i * n + j)] -= f * A[size_t(k * n + j)];
      for (int j = 0; j < ns; ++j) S[size_t(i * ns + j)] -= f * S[size_t(k * ns + j)];
    }
  }
  for (int j = 0; j < ns; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double v = S[size_t(i * ns + j)];
      for (int k = i + 1; k < n; ++k) v -= A[size_t(i * n + k)] * S[size_t(k * ns + j)];
      S[size_t(i * ns + j)] = v / A[size_t(i * n + i)];
    }
  return S;  // row = target index, column = source index
}

Space compile_space(const Grid& grid, const std::vector<int64_t>& active,
                    const std::vector<int32_t>& degrees, int basis, int cg_mask) {
  const int D = grid.D;
  if (basis != kFullTensor && basis != kTrunk) throw std::invalid_argument("unknown basis family");
  if (degrees.size() != active.size() * size_t(D))
    throw std::invalid_argument("leaf degree data must match the active-leaf count");
  Snapshot snap(grid, active);
  const size_t nl = snap.leaves.size();
  Space sp;
  sp.D = D;
  sp.leaves = snap.leaves;
  sp.boundary_leaf = snap.boundary_leaf;
  sp.boundary_axis = snap.boundary_axis;
  sp.boundary_side = snap.boundary_side;
  sp.iface_minus = snap.iface_minus;
  sp.iface_plus = snap.iface_plus;
  sp.iface_axis = snap.iface_axis;

  // Degrees in Morton order.
  std::unordered_map<int64_t, size_t> given;
  for (size_t i = 0; i < active.size(); ++i) given[active[i]] = i;
  sp.degrees.resize(nl);
  for (size_t li = 0; li < nl; ++li) {
    size_t src = given.at(snap.leaves[li]);
    std::array<int32_t, kMaxDim> d{{0, 0, 0}};
    for (int a = 0; a < D; ++a) {
      d[a] = degrees[src * size_t(D) + size_t(a)];
      if (d[a] < 0) throw std::invalid_argument("degrees must be non-negative");
      if (d[a] == 0 && (cg_mask >> a & 1))
        throw std::invalid_argument("degree zero is only allowed on DG axes");
    }
    sp.degrees[li] = d;
  }

  // Phase 1: patterns and lattice intervals.
  std::map<std::array<int32_t, kMaxDim>, Pattern> patterns;
  std::vector<const Pattern*> leaf_pattern(nl);
  std::array<int32_t, kMaxDim> finest{{0, 0, 0}};
  for (size_t li = 0; li < nl; ++li)
    for (int a = 0; a < D; ++a) finest[a] = std::max(finest[a], grid.level[snap.leaves[li]][a]);
  std::vector<std::array<int64_t, kMaxDim>> lower(nl), upper(nl);
  for (size_t li = 0; li < nl; ++li) {
    auto it = patterns.find(sp.degrees[li]);
    if (it == patterns.end())
      it = patterns.emplace(sp.degrees[li], make_pattern(basis, D, cg_mask, sp.degrees[li])).first;
    leaf_pattern[li] = &it->second;
    int64_t cell = snap.leaves[li];
    for (int a = 0; a < D; ++a) {
      int sh = finest[a] - grid.level[cell][a];
      lower[li][a] = grid.coord[cell][a] << sh;
      upper[li][a] = (grid.coord[cell][a] + 1) << sh;
    }
  }

  const bool fully_dg = (cg_mask & ((1 << D) - 1)) == 0;
  int64_t next_dof = 0;
  std::vector<int64_t> var_base(nl, 0);
  int64_t nvar = 0;
  std::vector<int64_t> roots;
  std::vector<int64_t> free_dof;
  std::vector<uint8_t> is_pivot;
  std::unordered_map<int64_t, std::vector<std::pair<int64_t, double>>> pivot_terms;

  if (!fully_dg) {
    // Phase 3: provisional boundary variables.
    for (size_t li = 0; li < nl; ++li) {
      var_base[li] = nvar;
      nvar += leaf_pattern[li]->coupled_count;
    }
    sp.boundary_variable_count = nvar;
    auto var_of = [&](size_t li, int32_t box) -> int64_t {
      int32_t r = leaf_pattern[li]->box_coupled_rank[size_t(box)];
      return r < 0 ? -1 : var_base[li] + r;
    };
    auto box_in = [&](size_t li, std::array<int32_t, kMaxDim> m) -> int32_t {
      const Pattern& p = *leaf_pattern[li];
      for (int a = 0; a < D; ++a)
        if (m[a] < 0 || m[a] > p.deg[a]) return -1;
      return flatten(m, p.shape, D);
    };

    // Phase 4: direct merges, then hanging / p-mismatch rows.
    UnionFind uf(nvar);
    std::vector<std::tuple<size_t, size_t, int>> constraint_pairs;
    for (size_t f = 0; f < snap.iface_minus.size(); ++f) {
      int axis = snap.iface_axis[f];
      if (!(cg_mask >> axis & 1)) continue;
      size_t first = size_t(snap.iface_minus[f]), second = size_t(snap.iface_plus[f]);
      bool matching = true;
      for (int a = 0; a < D && matching; ++a) {
        if (a == axis) continue;
        matching = lower[first][a] == lower[second][a] && upper[first][a] == upper[second][a];
      }
      const auto& fm = leaf_pattern[first]->face_modes[axis][1];
      const auto& sm = leaf_pattern[second]->face_modes[axis][0];
      bool same = matching && fm.size() == sm.size();
      if (same) {
        for (int32_t b : fm) {
          auto m = unflatten(b, leaf_pattern[first]->shape, D);
          m[axis] = 0;
          int32_t ob = box_in(second, m);
          if (ob < 0 || var_of(second, ob) < 0) {
            same = false;
            break;
          }
        }
      }
      if (same) {
        for (int32_t b : fm) {
          auto m = unflatten(b, leaf_pattern[first]->shape, D);
          m[axis] = 0;
          int64_t v2 = var_of(second, box_in(second, m));
          if (v2 < 0) throw std::invalid_argument("missing matching trace variable");
          uf.unite(var_of(first, b), v2);
        }
      } else {
        constraint_pairs.emplace_back(first, second, axis);
      }
    }
    roots.resize(size_t(nvar));
    for (int64_t v = 0; v < nvar; ++v) roots[size_t(v)] = uf.root(v);
    uf.parent.clear();
    uf.parent.shrink_to_fit();

    std::vector<Row> rows;
    std::map<std::tuple<int, int, int, int64_t>, std::vector<double>> rcache;
    for (const auto& cp : constraint_pairs) {
      size_t first = std::get<0>(cp), second = std::get<1>(cp);
      int axis = std::get<2>(cp);
      std::array<int64_t, kMaxDim> plo{{0, 0, 0}}, pup{{0, 0, 0}};
      std::array<int32_t, kMaxDim> cdeg{{0, 0, 0}};
      for (int a = 0; a < D; ++a) {
        if (a == axis) {
          plo[a] = pup[a] = upper[first][a];
          cdeg[a] = 0;
        } else {
          plo[a] = std::max(lower[first][a], lower[second][a]);
          pup[a] = std::min(upper[first][a], upper[second][a]);
          if (plo[a] >= pup[a])
            throw std::invalid_argument("adjacent leaves do not share a positive-measure face overlap");
          cdeg[a] = std::max(sp.degrees[first][a], sp.degrees[second][a]);
        }
      }
      auto coefficient = [&](size_t li, const std::array<int32_t, kMaxDim>& mode,
                             const std::array<int32_t, kMaxDim>& cmode) -> double {
        double c = 1.0;
        for (int a = 0; a < D; ++a) {
          if (a == axis) continue;
          int slev, tlev;
          int64_t scoord, tcoord;
          interval_level_and_coord(lower[li][a], upper[li][a], finest[a], slev, scoord);
          interval_level_and_coord(plo[a], pup[a], finest[a], tlev, tcoord);
          int delta = tlev - slev;
          if (delta < 0) throw std::invalid_argument("target interval must not be coarser than source interval");
          int64_t rel = tcoord - (scoord << delta);
          if (rel < 0 || rel >= (int64_t(1) << delta))
            throw std::invalid_argument("target interval must be contained in source interval");
          auto key = std::make_tuple(int(sp.degrees[li][a]), int(cdeg[a]), delta, rel);
          auto it = rcache.find(key);
          if (it == rcache.end())
            it = rcache.emplace(key, affine_restriction_matrix(sp.degrees[li][a], cdeg[a], delta, rel)).first;
          int ns = sp.degrees[li][a] + 1;
          c *= it->second[size_t(cmode[a] * ns + mode[a])];
          if (c == 0.0) return 0.0;
        }
        return c;
      };
      auto cmodes = basis_modes(kFullTensor, D, cdeg.data());
      for (const auto& cm : cmodes) {
        Row row;
        for (int32_t b : leaf_pattern[first]->face_modes[axis][1]) {
          auto m = unflatten(b, leaf_pattern[first]->shape, D);
          double c = coefficient(first, m, cm);
          if (c == 0.0) continue;
          row.add(roots[size_t(var_of(first, b))], c);
        }
        for (int32_t b : leaf_pattern[second]->face_modes[axis][0]) {
          auto m = unflatten(b, leaf_pattern[second]->shape, D);
          double c = coefficient(second, m, cm);
          if (c == 0.0) continue;
          row.add(roots[size_t(var_of(second, b))], -c);
        }
        row.cleanup();
        if (!row.empty()) rows.push_back(std::move(row));
      }
    }
    sp.constraint_row_count = int64_t(rows.size());

    // Phase 5: component-wise elimination.
    is_pivot.assign(size_t(nvar), 0);
    std::vector<std::vector<int64_t>> component_orders;
    std::unordered_map<int64_t, Row> pivots;
    if (!rows.empty()) {
      UnionFind cu(nvar);
      for (const auto& r : rows)
        for (size_t i = 1; i < r.var.size(); ++i) cu.unite(r.var[0], r.var[i]);
      std::map<int64_t, std::vector<size_t>> components;
      for (size_t i = 0; i < rows.size(); ++i) components[cu.root(rows[i].var[0])].push_back(i);
      for (auto& comp : components) {
        std::unordered_map<int64_t, const Row*> local;
        std::vector<int64_t> order;
        for (size_t ri : comp.second) {
          Row& row = rows[ri];
          while (true) {
            bool changed = false;
            for (size_t i = 0; i < row.var.size(); ++i) {
              auto it = local.find(row.var[i]);
              if (it == local.end()) continue;
              double factor = row.coef[i];
              row.remove_at(i);
              const Row& ex = *it->second;
              for (size_t j = 0; j < ex.var.size(); ++j) row.add(ex.var[j], factor * ex.coef[j]);
              changed = true;
              break;
            }
            row.cleanup();
            if (!changed) break;
          }
          if (row.empty()) continue;
          size_t pi = 0;
          for (size_t i = 1; i < row.var.size(); ++i)
            if (row.var[i] > row.var[pi]) pi = i;
          int64_t pv = row.var[pi];
          double pc = row.coef[pi];
          row.remove_at(pi);
          double scale = -1.0 / pc;
          Row ex;
          for (size_t j = 0; j < row.var.size(); ++j) ex.add(row.var[j], scale * row.coef[j]);
          ex.cleanup();
          auto ins = pivots.emplace(pv, std::move(ex));
          if (!ins.second) throw std::runtime_error("boundary variable eliminated twice");
          local[pv] = &ins.first->second;
          is_pivot[size_t(pv)] = 1;
          order.push_back(pv);
        }
        component_orders.push_back(std::move(order));
      }
    }

    // Free boundary dofs are numbered first, in variable order.
    free_dof.assign(size_t(nvar), -1);
    for (int64_t v = 0; v < nvar; ++v)
      if (roots[size_t(v)] == v && !is_pivot[size_t(v)]) free_dof[size_t(v)] = next_dof++;

    for (const auto& order : component_orders)
      for (auto it = order.rbegin(); it != order.rend(); ++it) {
        const Row& ex = pivots.at(*it);
        Row acc;
        for (size_t j = 0; j < ex.var.size(); ++j) {
          int64_t v = ex.var[j];
          double factor = ex.coef[j];
          if (factor == 0.0) continue;
          if (is_pivot[size_t(v)]) {
            auto pt = pivot_terms.find(v);
            if (pt == pivot_terms.end()) throw std::runtime_error("pivot expansion used before it was finalized");
            for (const auto& t : pt->second) acc.add(t.first, factor * t.second);
          } else {
            if (free_dof[size_t(v)] < 0) throw std::runtime_error("boundary term expansion references a non-root variable");
            acc.add(free_dof[size_t(v)], factor);
          }
        }
        acc.cleanup();
        std::vector<std::pair<int64_t, double>> sorted;
        for (size_t j = 0; j < acc.var.size(); ++j) sorted.emplace_back(acc.var[j], acc.coef[j]);
        std::sort(sorted.begin(), sorted.end(),
                  [](const std::pair<int64_t, double>& x, const std::pair<int64_t, double>& y) {
                    return x.first < y.first;
                  });
        pivot_terms[*it] = std::move(sorted);
      }
  }

  // Phase 6: per-leaf expansions.
  sp.mode_offsets.assign(nl + 1, 0);
  sp.term_offsets.push_back(0);
  for (size_t li = 0; li < nl; ++li) {
    const Pattern& p = *leaf_pattern[li];
    for (size_t k = 0; k < p.modes.size(); ++k) {
      const auto& m = p.modes[k];
      size_t before = sp.term_indices.size();
      if (!fully_dg && requires_coupling(D, cg_mask, m)) {
        int32_t r = p.box_coupled_rank[size_t(p.mode_box[k])];
        if (r < 0) throw std::invalid_argument("missing boundary variable for local mode");
        int64_t root = roots[size_t(var_base[li] + r)];
        if (is_pivot[size_t(root)]) {
          for (const auto& t : pivot_terms.at(root)) {
            if (t.second == 0.0) continue;
            sp.term_indices.push_back(t.first);
            sp.term_coefficients.push_back(t.second);
          }
        } else {
          sp.term_indices.push_back(free_dof[size_t(root)]);
          sp.term_coefficients.push_back(1.0);
        }
      } else {
        sp.term_indices.push_back(next_dof++);
        sp.term_coefficients.push_back(1.0);
      }
      if (sp.term_indices.size() == before) continue;  // constrained to zero: omitted
      sp.mode_box.push_back(p.mode_box[k]);
      sp.term_offsets.push_back(int64_t(sp.term_indices.size()));
    }
    sp.mode_offsets[li + 1] = int64_t(sp.mode_box.size());
  }
  sp.dof_count = next_dof;
  return sp;
}

}  // namespace gh
