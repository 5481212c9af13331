// Compiled hp space: per-leaf sparse expansions of local tensor-product modes in
// global scalar dofs.
//
// Follows src/continuity.jl (phases 26-38): provisional boundary variables per
// leaf in local-mode order (222-250), union-find merge of matching faces keeping
// the smaller id (466-480, 1184-1198), hanging / p-mismatch comparison rows
// (412-446, 509-521, 1220-1300), component-wise elimination with the largest
// variable id as pivot (757-800), free boundary dofs numbered first in variable
// order (870-880), then interior modes leaf by leaf (252-350), and the dyadic
// affine restriction matrices (1158-1198). Mode order is mixed radix with axis 1
// fastest (src/basis.jl:150-175).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "topology.hpp"

namespace gh {

enum BasisKind { kFullTensor = 0, kTrunk = 1 };

struct Space {
  int D = 0;
  int64_t dof_count = 0;
  std::vector<int64_t> leaves;                       // Morton order
  std::vector<std::array<int32_t, kMaxDim>> degrees;  // per leaf
  // Per leaf: active local modes (flattened box index, axis 1 fastest).
  std::vector<int64_t> mode_offsets;  // size n_leaves + 1
  std::vector<int32_t> mode_box;      // flattened box index per active mode
  // Per active mode: CSR expansion in global dofs (0-based).
  std::vector<int64_t> term_offsets;  // size n_modes + 1
  std::vector<int64_t> term_indices;
  std::vector<double> term_coefficients;
  // Snapshot data (leaf indices are positions in `leaves`).
  std::vector<int64_t> boundary_leaf;
  std::vector<int32_t> boundary_axis, boundary_side;
  std::vector<int64_t> iface_minus, iface_plus;
  std::vector<int32_t> iface_axis;
  int64_t boundary_variable_count = 0;
  int64_t constraint_row_count = 0;
};

// Compile a space on the active set `active` of `grid`. `degrees` holds D
// degrees per entry of `active` (same order). `cg_mask` bit a set means axis a
// is continuous.
Space compile_space(const Grid& grid, const std::vector<int64_t>& active,
                    const std::vector<int32_t>& degrees, int basis, int cg_mask);

// Local mode list (mixed radix, axis 1 fastest) of one degree tuple.
std::vector<std::array<int32_t, kMaxDim>> basis_modes(int basis, int D, const int32_t* degrees);

// Exact restriction of a source FE basis on an interval to a target basis on a
// dyadic subinterval; returns (target_degree+1) x (source_degree+1) row-major.
std::vector<double> affine_restriction_matrix(int source_degree, int target_degree, int delta,
                                              int64_t relative);

}  // namespace gh
