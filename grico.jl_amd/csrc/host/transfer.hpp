// Reduced-space prolongation between two nested hp spaces of one grid: the
// exact embedding of every coarse local mode in the fine modal basis (tensor
// product of one-dimensional dyadic restriction matrices), composed with both
// reduced maps.
//
// Follows src/multigrid.jl:614-857: _compile_multigrid_transfer (614-636),
// _append_field_reduced_transfer_entries! (638-692), _compile_local_transfer_entries
// (694-717), _fine_leaves_under_coarse_leaf (724-744), _axis_transfer_matrix!
// (746-757), _local_mode_reduced_expansion (782-802), _fine_mode_reduced_coordinate
// (804-821) and the duplicate-entry consistency check (829-852).
#pragma once
#include <cstdint>
#include <vector>

#include "topology.hpp"

namespace gh {

// Borrowed view of a compiled space plus the reduced map of a plan on it.
struct SpaceView {
  int D = 0;
  int64_t leaf_count = 0;
  const int64_t* leaves = nullptr;        // cell ids
  const int32_t* degrees = nullptr;       // leaf_count x D
  const int64_t* mode_offsets = nullptr;  // leaf_count + 1
  const int32_t* mode_box = nullptr;
  const int64_t* term_offsets = nullptr;  // modes + 1
  const int64_t* term_indices = nullptr;
  const double* term_coefficients = nullptr;
  int64_t full_dof_count = 0;
  const int64_t* map_offsets = nullptr;   // full_dof_count + 1 (full -> reduced CSR)
  const int64_t* map_indices = nullptr;
  const double* map_coefficients = nullptr;
  const int64_t* reduced_index = nullptr; // full dof -> reduced dof, -1 when not a solve dof
};

struct Transfer {
  // Sorted by (fine, coarse).
  std::vector<int64_t> coarse;
  std::vector<int64_t> fine;
  std::vector<double> coefficient;
};

Transfer build_transfer(const Grid& grid, const SpaceView& coarse, const SpaceView& fine);

}  // namespace gh
