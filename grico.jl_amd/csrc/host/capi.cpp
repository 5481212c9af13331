// C ABI of the host plan compiler (include/grico_host.h).
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../../include/grico_host.h"
#include "polys.hpp"
#include "space.hpp"
#include "topology.hpp"
#include "transfer.hpp"

struct gh_grid {
  gh::Grid g;
};
struct gh_space {
  gh::Space s;
};
struct gh_transfer {
  gh::Transfer t;
};

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return -1;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -2;
  }
}
}  // namespace

extern "C" {

const char* gh_last_error(void) { return g_error.c_str(); }

int gh_gauss_legendre(int point_count, double* points, double* weights) {
  return guarded([&] { gh::gauss_legendre(point_count, points, weights); });
}

int gh_fe_basis(double x, int degree, double* values, double* derivatives) {
  return guarded([&] {
    if (degree < 0) throw std::invalid_argument("degree must be non-negative");
    gh::fe_basis(x, degree, values, derivatives);
  });
}

int gh_restriction(int source_degree, int target_degree, int delta, int64_t relative, double* out) {
  return guarded([&] {
    auto m = gh::affine_restriction_matrix(source_degree, target_degree, delta, relative);
    std::memcpy(out, m.data(), m.size() * sizeof(double));
  });
}

int gh_grid_create(int dim, const int64_t* root_counts, gh_grid** out) {
  return guarded([&] { *out = new gh_grid{gh::Grid(dim, root_counts)}; });
}

void gh_grid_destroy(gh_grid* grid) { delete grid; }

int64_t gh_grid_cell_count(const gh_grid* grid) { return grid->g.cell_count(); }

int gh_grid_refine(gh_grid* grid, int64_t cell, int axis, int64_t* first_child) {
  return guarded([&] { *first_child = grid->g.refine(cell, axis); });
}

int gh_grid_cell(const gh_grid* grid, int64_t cell, int32_t* levels, int64_t* coords, int64_t* parent,
                 int64_t* first_child, int32_t* split_axis, int32_t* active) {
  return guarded([&] {
    const gh::Grid& g = grid->g;
    if (cell < 0 || cell >= g.cell_count()) throw std::invalid_argument("cell out of range");
    for (int a = 0; a < g.D; ++a) {
      if (levels) levels[a] = g.level[size_t(cell)][size_t(a)];
      if (coords) coords[a] = g.coord[size_t(cell)][size_t(a)];
    }
    if (parent) *parent = g.parent[size_t(cell)];
    if (first_child) *first_child = g.first_child[size_t(cell)];
    if (split_axis) *split_axis = g.split[size_t(cell)];
    if (active) *active = g.active[size_t(cell)];
  });
}

int gh_grid_export(const gh_grid* grid, int32_t* levels, int64_t* coords, int64_t* parent,
                   int64_t* first_child, int32_t* split_axis, uint8_t* active) {
  return guarded([&] {
    const gh::Grid& g = grid->g;
    for (int64_t c = 0; c < g.cell_count(); ++c) {
      size_t i = size_t(c);
      for (int a = 0; a < g.D; ++a) {
        if (levels) levels[i * size_t(g.D) + size_t(a)] = g.level[i][size_t(a)];
        if (coords) coords[i * size_t(g.D) + size_t(a)] = g.coord[i][size_t(a)];
      }
      if (parent) parent[i] = g.parent[i];
      if (first_child) first_child[i] = g.first_child[i];
      if (split_axis) split_axis[i] = g.split[i];
      if (active) active[i] = g.active[i];
    }
  });
}

int gh_space_compile(const gh_grid* grid, int64_t leaf_count, const int64_t* leaves,
                     const int32_t* degrees, int basis, int cg_mask, gh_space** out) {
  return guarded([&] {
    std::vector<int64_t> act(leaves, leaves + leaf_count);
    std::vector<int32_t> deg(degrees, degrees + leaf_count * grid->g.D);
    auto sp = std::make_unique<gh_space>();
    sp->s = gh::compile_space(grid->g, act, deg, basis, cg_mask);
    *out = sp.release();
  });
}

void gh_space_destroy(gh_space* space) { delete space; }

int gh_space_sizes_get(const gh_space* space, gh_space_sizes* sz) {
  return guarded([&] {
    const gh::Space& s = space->s;
    sz->leaf_count = int64_t(s.leaves.size());
    sz->dof_count = s.dof_count;
    sz->mode_count = int64_t(s.mode_box.size());
    sz->term_count = int64_t(s.term_indices.size());
    sz->boundary_count = int64_t(s.boundary_leaf.size());
    sz->interface_count = int64_t(s.iface_minus.size());
    sz->boundary_variable_count = s.boundary_variable_count;
    sz->constraint_row_count = s.constraint_row_count;
  });
}

int gh_space_export(const gh_space* space, int64_t* leaves, int32_t* degrees, int64_t* mode_offsets,
                    int32_t* mode_box, int64_t* term_offsets, int64_t* term_indices,
                    double* term_coefficients) {
  return guarded([&] {
    const gh::Space& s = space->s;
    if (leaves) std::memcpy(leaves, s.leaves.data(), s.leaves.size() * sizeof(int64_t));
    if (degrees)
      for (size_t i = 0; i < s.degrees.size(); ++i)
        for (int a = 0; a < s.D; ++a) degrees[i * size_t(s.D) + size_t(a)] = s.degrees[i][size_t(a)];
    if (mode_offsets) std::memcpy(mode_offsets, s.mode_offsets.data(), s.mode_offsets.size() * sizeof(int64_t));
    if (mode_box) std::memcpy(mode_box, s.mode_box.data(), s.mode_box.size() * sizeof(int32_t));
    if (term_offsets) std::memcpy(term_offsets, s.term_offsets.data(), s.term_offsets.size() * sizeof(int64_t));
    if (term_indices) std::memcpy(term_indices, s.term_indices.data(), s.term_indices.size() * sizeof(int64_t));
    if (term_coefficients)
      std::memcpy(term_coefficients, s.term_coefficients.data(), s.term_coefficients.size() * sizeof(double));
  });
}

int gh_space_faces(const gh_space* space, int64_t* boundary_leaf, int32_t* boundary_axis,
                   int32_t* boundary_side, int64_t* interface_minus, int64_t* interface_plus,
                   int32_t* interface_axis) {
  return guarded([&] {
    const gh::Space& s = space->s;
    size_t nb = s.boundary_leaf.size(), ni = s.iface_minus.size();
    if (boundary_leaf) std::memcpy(boundary_leaf, s.boundary_leaf.data(), nb * sizeof(int64_t));
    if (boundary_axis) std::memcpy(boundary_axis, s.boundary_axis.data(), nb * sizeof(int32_t));
    if (boundary_side) std::memcpy(boundary_side, s.boundary_side.data(), nb * sizeof(int32_t));
    if (interface_minus) std::memcpy(interface_minus, s.iface_minus.data(), ni * sizeof(int64_t));
    if (interface_plus) std::memcpy(interface_plus, s.iface_plus.data(), ni * sizeof(int64_t));
    if (interface_axis) std::memcpy(interface_axis, s.iface_axis.data(), ni * sizeof(int32_t));
  });
}


int gh_transfer_create(const gh_grid* grid, const gh_space_view* coarse, const gh_space_view* fine,
                       gh_transfer** out) {
  return guarded([&] {
    if (!grid || !coarse || !fine || !out) throw std::invalid_argument("null argument");
    auto view = [](const gh_space_view* v) {
      gh::SpaceView s;
      s.D = v->dim;
      s.leaf_count = v->leaf_count;
      s.leaves = v->leaves;
      s.degrees = v->degrees;
      s.mode_offsets = v->mode_offsets;
      s.mode_box = v->mode_box;
      s.term_offsets = v->term_offsets;
      s.term_indices = v->term_indices;
      s.term_coefficients = v->term_coefficients;
      s.full_dof_count = v->full_dof_count;
      s.map_offsets = v->map_offsets;
      s.map_indices = v->map_indices;
      s.map_coefficients = v->map_coefficients;
      s.reduced_index = v->reduced_index;
      if (!s.leaves || !s.degrees || !s.mode_offsets || !s.mode_box || !s.term_offsets || !s.term_indices ||
          !s.term_coefficients || !s.map_offsets || !s.map_indices || !s.map_coefficients || !s.reduced_index)
        throw std::invalid_argument("incomplete space view");
      return s;
    };
    auto t = std::make_unique<gh_transfer>();
    t->t = gh::build_transfer(grid->g, view(coarse), view(fine));
    *out = t.release();
  });
}

void gh_transfer_destroy(gh_transfer* transfer) { delete transfer; }

int64_t gh_transfer_nnz(const gh_transfer* transfer) { return int64_t(transfer->t.fine.size()); }

int gh_transfer_export(const gh_transfer* transfer, int64_t* coarse_indices, int64_t* fine_indices,
                       double* coefficients) {
  return guarded([&] {
    const gh::Transfer& t = transfer->t;
    const size_t n = t.fine.size();
    if (coarse_indices) std::memcpy(coarse_indices, t.coarse.data(), n * sizeof(int64_t));
    if (fine_indices) std::memcpy(fine_indices, t.fine.data(), n * sizeof(int64_t));
    if (coefficients) std::memcpy(coefficients, t.coefficient.data(), n * sizeof(double));
  });
}

int gh_conflict_coloring(int64_t block_count, const int64_t* block_ptr, const int64_t* block_dofs,
                         int64_t item_count, const int64_t* item_ptr, const int64_t* item_dofs, int64_t dof_count,
                         int32_t* colors, int32_t* color_count) {
  return guarded([&] {
    if (block_count < 0 || item_count < 0 || dof_count < 0 || !colors || !color_count ||
        (block_count > 0 && (!block_ptr || !block_dofs)) || (item_count > 0 && (!item_ptr || !item_dofs)))
      throw std::invalid_argument("invalid coloring arguments");
    // dof -> items touching it (CSR)
    std::vector<int64_t> dptr(size_t(dof_count) + 1, 0);
    for (int64_t i = 0; i < item_count; ++i)
      for (int64_t t = item_ptr[i]; t < item_ptr[i + 1]; ++t) {
        const int64_t d = item_dofs[t];
        if (d < 0 || d >= dof_count) throw std::invalid_argument("item dof out of range");
        ++dptr[size_t(d) + 1];
      }
    for (int64_t d = 0; d < dof_count; ++d) dptr[size_t(d) + 1] += dptr[size_t(d)];
    std::vector<int64_t> ditems(size_t(dptr.back())), fillp(dptr.begin(), dptr.end() - 1);
    for (int64_t i = 0; i < item_count; ++i)
      for (int64_t t = item_ptr[i]; t < item_ptr[i + 1]; ++t) ditems[size_t(fillp[size_t(item_dofs[t])]++)] = i;
    // item -> blocks colored so far that touch it
    std::vector<std::vector<int32_t>> item_blocks(static_cast<size_t>(item_count));
    std::vector<int64_t> stamp(static_cast<size_t>(item_count), -1);
    std::vector<int64_t> used;  // color -> last block that saw it forbidden
    int32_t ncol = 0;
    std::vector<int64_t> touched;
    for (int64_t b = 0; b < block_count; ++b) {
      touched.clear();
      for (int64_t t = block_ptr[b]; t < block_ptr[b + 1]; ++t) {
        const int64_t d = block_dofs[t];
        if (d < 0 || d >= dof_count) throw std::invalid_argument("block dof out of range");
        for (int64_t k = dptr[size_t(d)]; k < dptr[size_t(d) + 1]; ++k) {
          const int64_t i = ditems[size_t(k)];
          if (stamp[size_t(i)] != b) {
            stamp[size_t(i)] = b;
            touched.push_back(i);
          }
        }
      }
      for (int64_t i : touched)
        for (int32_t ob : item_blocks[size_t(i)]) used[size_t(colors[ob])] = b;
      int32_t c = 0;
      while (c < ncol && used[size_t(c)] == b) ++c;
      if (c == ncol) {
        ++ncol;
        used.push_back(-1);
      }
      colors[b] = c;
      for (int64_t i : touched) item_blocks[size_t(i)].push_back(int32_t(b));
    }
    *color_count = ncol;
  });
}

}  // extern "C"
