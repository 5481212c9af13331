// Reduced-space multigrid transfers: see transfer.hpp for the reference citations.
#include "transfer.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <tuple>

#include "space.hpp"

namespace gh {
namespace {

struct Triple {
  int64_t fine, coarse;
  double value;
};

// Sparse accumulator keyed by reduced dof, kept in first-insertion order.
struct Accum {
  std::vector<int64_t> key;
  std::vector<double> val;
  void add(int64_t k, double v) {
    if (v == 0.0) return;
    for (size_t i = 0; i < key.size(); ++i)
      if (key[i] == k) {
        val[i] += v;
        return;
      }
    key.push_back(k);
    val.push_back(v);
  }
  void clear() {
    key.clear();
    val.clear();
  }
};

std::vector<int64_t> position_of(const Grid& grid, const SpaceView& s) {
  std::vector<int64_t> pos(size_t(grid.cell_count()), -1);
  for (int64_t i = 0; i < s.leaf_count; ++i) {
    int64_t c = s.leaves[i];
    if (c < 0 || c >= grid.cell_count()) throw std::invalid_argument("space leaf is not a grid cell");
    pos[size_t(c)] = i;
  }
  return pos;
}

std::array<int32_t, kMaxDim> box_shape(const SpaceView& s, int64_t li) {
  std::array<int32_t, kMaxDim> sh{{1, 1, 1}};
  for (int a = 0; a < s.D; ++a) sh[a] = s.degrees[li * s.D + a] + 1;
  return sh;
}

}  // namespace

Transfer build_transfer(const Grid& grid, const SpaceView& coarse, const SpaceView& fine) {
  const int D = grid.D;
  if (coarse.D != D || fine.D != D) throw std::invalid_argument("transfer spaces must match the grid dimension");
  const std::vector<int64_t> fine_pos = position_of(grid, fine);

  std::map<std::tuple<int, int, int, int64_t>, std::vector<double>> cache;
  auto matrix = [&](int sdeg, int tdeg, int delta, int64_t rel) -> const std::vector<double>& {
    auto key = std::make_tuple(sdeg, tdeg, delta, rel);
    auto it = cache.find(key);
    if (it == cache.end()) {
      // The collocation solve leaves rounding residue (~1e-16) where the
      // embedding is exactly zero (a p-embedding of the hierarchical basis is
      // an identity block); such residue would turn every fine mode into a
      // sum over all coarse modes. Entries below 4 eps of the matrix scale are
      // dropped, the tolerance the continuity rows use for cancellation.
      std::vector<double> m = affine_restriction_matrix(sdeg, tdeg, delta, rel);
      double scale = 1.0;
      for (double v : m) scale = std::max(scale, std::fabs(v));
      const double tol = 4.0 * std::numeric_limits<double>::epsilon() * scale;
      for (double& v : m)
        if (std::fabs(v) <= tol) v = 0.0;
      it = cache.emplace(key, std::move(m)).first;
    }
    return it->second;
  };

  std::vector<Triple> triples;
  std::vector<Accum> coarse_exp;  // per coarse local mode: reduced expansion
  std::vector<Accum> target;      // per fine local mode
  std::vector<int32_t> fine_lookup;

  for (int64_t ci = 0; ci < coarse.leaf_count; ++ci) {
    const int64_t cc = coarse.leaves[ci];
    // Fine leaves covering this coarse leaf: itself (p-level) or its two
    // children (h-level).
    int64_t fl[2];
    int nfl = 0;
    if (fine_pos[size_t(cc)] >= 0) {
      fl[nfl++] = fine_pos[size_t(cc)];
    } else {
      const int64_t first = grid.first_child[size_t(cc)];
      if (first < 0)
        throw std::invalid_argument("fine multigrid space does not contain a coarse leaf or its immediate children");
      if (fine_pos[size_t(first)] < 0 || fine_pos[size_t(first + 1)] < 0)
        throw std::invalid_argument("h-multigrid transfer requires adjacent dyadic levels");
      fl[nfl++] = fine_pos[size_t(first)];
      fl[nfl++] = fine_pos[size_t(first + 1)];
    }

    // Reduced expansion of every coarse local mode (src/multigrid.jl:782-802).
    const auto cshape = box_shape(coarse, ci);
    const int64_t m0 = coarse.mode_offsets[ci], m1 = coarse.mode_offsets[ci + 1];
    coarse_exp.assign(size_t(m1 - m0), Accum());
    for (int64_t m = m0; m < m1; ++m) {
      Accum& acc = coarse_exp[size_t(m - m0)];
      for (int64_t t = coarse.term_offsets[m]; t < coarse.term_offsets[m + 1]; ++t) {
        const int64_t dof = coarse.term_indices[t];
        const double c = coarse.term_coefficients[t];
        for (int64_t k = coarse.map_offsets[dof]; k < coarse.map_offsets[dof + 1]; ++k)
          acc.add(coarse.map_indices[k], c * coarse.map_coefficients[k]);
      }
    }

    for (int f = 0; f < nfl; ++f) {
      const int64_t fi = fl[f];
      const int64_t fc = fine.leaves[fi];
      const auto fshape = box_shape(fine, fi);
      for (int a = 0; a < D; ++a)
        if (fshape[a] > 17) throw std::invalid_argument("multigrid transfer supports degrees up to 16");
      const std::vector<double>* R[kMaxDim] = {nullptr, nullptr, nullptr};
      for (int a = 0; a < D; ++a) {
        const int sdeg = cshape[a] - 1, tdeg = fshape[a] - 1;
        if (tdeg < sdeg) throw std::invalid_argument("fine multigrid level must represent every coarse mode exactly");
        const int delta = grid.level[size_t(fc)][a] - grid.level[size_t(cc)][a];
        if (delta < 0) throw std::invalid_argument("fine multigrid leaf is coarser than its coarse leaf");
        const int64_t rel = grid.coord[size_t(fc)][a] - (grid.coord[size_t(cc)][a] << delta);
        if (rel < 0 || rel >= (int64_t(1) << delta))
          throw std::invalid_argument("fine multigrid leaf is not contained in its coarse leaf");
        R[a] = &matrix(sdeg, tdeg, delta, rel);
      }
      // fine box index -> local mode position
      int64_t fbox = 1;
      for (int a = 0; a < D; ++a) fbox *= fshape[a];
      fine_lookup.assign(size_t(fbox), -1);
      const int64_t f0 = fine.mode_offsets[fi], f1 = fine.mode_offsets[fi + 1];
      for (int64_t m = f0; m < f1; ++m) fine_lookup[size_t(fine.mode_box[m])] = int32_t(m - f0);
      target.assign(size_t(f1 - f0), Accum());

      // Local entries in coarse-mode order (src/multigrid.jl:694-717); every
      // nonzero tensor coefficient feeds the fine mode's reduced expansion.
      for (int64_t m = m0; m < m1; ++m) {
        const Accum& src = coarse_exp[size_t(m - m0)];
        if (src.key.empty()) continue;
        int32_t cb = coarse.mode_box[m];
        int cm[kMaxDim] = {0, 0, 0};
        for (int a = 0; a < D; ++a) {
          cm[a] = cb % cshape[a];
          cb /= cshape[a];
        }
        // Enumerate the fine tensor indices whose per-axis factors are all
        // nonzero (p-embeddings are identities, h-embeddings triangular).
        int nz[kMaxDim][17];
        int nnz[kMaxDim] = {1, 1, 1};
        for (int a = 0; a < D; ++a) {
          nnz[a] = 0;
          for (int fidx = 0; fidx < fshape[a]; ++fidx)
            if ((*R[a])[size_t(fidx * cshape[a] + cm[a])] != 0.0) nz[a][nnz[a]++] = fidx;
          if (nnz[a] == 0) break;
        }
        bool empty = false;
        for (int a = 0; a < D; ++a) empty = empty || nnz[a] == 0;
        if (empty) continue;
        int pos[kMaxDim] = {0, 0, 0};
        while (true) {
          double coef = 1.0;
          int64_t box = 0, stride = 1;
          for (int a = 0; a < D; ++a) {
            const int fidx = nz[a][pos[a]];
            coef *= (*R[a])[size_t(fidx * cshape[a] + cm[a])];
            box += fidx * stride;
            stride *= fshape[a];
          }
          if (coef != 0.0) {
            if (!std::isfinite(coef)) throw std::invalid_argument("multigrid transfer coefficient must be finite");
            const int32_t lm = fine_lookup[size_t(box)];
            if (lm >= 0) {
              Accum& dst = target[size_t(lm)];
              for (size_t k = 0; k < src.key.size(); ++k) dst.add(src.key[k], coef * src.val[k]);
            }
          }
          int a = 0;
          for (; a < D; ++a) {
            if (++pos[a] < nnz[a]) break;
            pos[a] = 0;
          }
          if (a == D) break;
        }
      }

      // Fine coordinates: single-term modes on solve dofs only
      // (src/multigrid.jl:804-821).
      for (int64_t m = f0; m < f1; ++m) {
        const Accum& acc = target[size_t(m - f0)];
        if (acc.key.empty()) continue;
        const int64_t t0 = fine.term_offsets[m], t1 = fine.term_offsets[m + 1];
        if (t1 != t0 + 1) continue;
        const int64_t dof = fine.term_indices[t0];
        const int64_t red = fine.reduced_index[dof];
        if (red < 0) continue;
        const double inv = 1.0 / fine.term_coefficients[t0];
        for (size_t k = 0; k < acc.key.size(); ++k) {
          const double v = inv * acc.val[k];
          if (!std::isfinite(v)) throw std::invalid_argument("multigrid transfer coefficient must be finite");
          if (v == 0.0) continue;
          triples.push_back({red, acc.key[k], v});
        }
      }
    }
  }

  // Deduplicate shared fine dofs: every occurrence must agree
  // (src/multigrid.jl:824-852); output sorted by (fine, coarse).
  std::stable_sort(triples.begin(), triples.end(), [](const Triple& x, const Triple& y) {
    return x.fine != y.fine ? x.fine < y.fine : x.coarse < y.coarse;
  });
  Transfer out;
  const double eps = std::numeric_limits<double>::epsilon();
  for (size_t i = 0; i < triples.size();) {
    size_t j = i + 1;
    const double first = triples[i].value;
    while (j < triples.size() && triples[j].fine == triples[i].fine && triples[j].coarse == triples[i].coarse) {
      const double second = triples[j].value;
      const double scale = std::max(1.0, std::max(std::fabs(first), std::fabs(second)));
      if (std::fabs(first - second) > 10000.0 * eps * scale)
        throw std::invalid_argument("inconsistent duplicate multigrid transfer entry");
      ++j;
    }
    out.fine.push_back(triples[i].fine);
    out.coarse.push_back(triples[i].coarse);
    out.coefficient.push_back(first);
    i = j;
  }
  return out;
}

}  // namespace gh
