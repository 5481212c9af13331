/*
 * grico_host.h — host-side plan compiler (CPU) for the MI355X matrix-free
 * backend.
 *
 * These entry points restate the setup half of Grico's compile(problem) that a
 * device backend needs: the refinement tree, the Morton-ordered active-leaf
 * snapshot and the continuity compiler that turns local tensor-product modes
 * into sparse expansions in global scalar dofs. In a Julia deployment these
 * tables come from Grico itself (HpSpace / CellValues term maps) and only the
 * device entry points in grico_mf.h are bound; this library exists so the
 * Python harness in grico.jl_amd/gricoamd can compile the same tables without
 * Julia.
 *
 * Reference counterparts:
 *   gh_grid_*            src/topology.jl:96-130 (CartesianGrid), src/refinement.jl:47 (refine!)
 *   gh_space_compile     src/space.jl:300-326 (HpSpace) -> src/continuity.jl:153-422
 *   gh_space_faces       src/topology.jl:1502-1545 (boundary faces, interfaces of the snapshot)
 *   gh_gauss_legendre    src/quadrature.jl:342-365 (gauss_legendre_rule)
 *   gh_fe_basis          src/polynomials.jl:238-247, 403-434
 *   gh_restriction       src/continuity.jl:1158-1198 (_affine_restriction_matrix)
 *   gh_transfer_*        src/multigrid.jl:614-857 (_compile_multigrid_transfer)
 *
 * Conventions: all indices are 0-based; axes are 0-based; side 0 = lower,
 * side 1 = upper. Functions return 0 on success and a negative status on
 * failure; gh_last_error() describes the most recent failure on this thread.
 */
#ifndef GRICO_HOST_H
#define GRICO_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gh_grid gh_grid;
typedef struct gh_space gh_space;

const char* gh_last_error(void);

/* Polynomials and quadrature. */
int gh_gauss_legendre(int point_count, double* points, double* weights);
int gh_fe_basis(double x, int degree, double* values, double* derivatives);
int gh_restriction(int source_degree, int target_degree, int delta, int64_t relative,
                   double* out /* (target_degree+1) x (source_degree+1), row-major */);

/* Refinement tree. */
int gh_grid_create(int dim, const int64_t* root_counts, gh_grid** out);
void gh_grid_destroy(gh_grid* grid);
int64_t gh_grid_cell_count(const gh_grid* grid);
/* Split an active leaf normal to `axis`; writes the first child id. */
int gh_grid_refine(gh_grid* grid, int64_t cell, int axis, int64_t* first_child);
/* levels/coords: dim entries each. */
int gh_grid_cell(const gh_grid* grid, int64_t cell, int32_t* levels, int64_t* coords,
                 int64_t* parent, int64_t* first_child, int32_t* split_axis, int32_t* active);
/* Bulk export of every stored cell (levels/coords: cell_count*dim). */
int gh_grid_export(const gh_grid* grid, int32_t* levels, int64_t* coords, int64_t* parent,
                   int64_t* first_child, int32_t* split_axis, uint8_t* active);

/* Continuity compile on an active-leaf set. degrees: dim entries per leaf, in
 * the order of `leaves`. basis: 0 full tensor, 1 trunk. cg_mask: bit a set
 * means axis a is continuous. */
int gh_space_compile(const gh_grid* grid, int64_t leaf_count, const int64_t* leaves,
                     const int32_t* degrees, int basis, int cg_mask, gh_space** out);
void gh_space_destroy(gh_space* space);

typedef struct gh_space_sizes {
  int64_t leaf_count;
  int64_t dof_count;
  int64_t mode_count;      /* active local modes summed over leaves */
  int64_t term_count;      /* expansion terms summed over modes */
  int64_t boundary_count;  /* boundary faces */
  int64_t interface_count; /* interface pairs */
  int64_t boundary_variable_count;
  int64_t constraint_row_count;
} gh_space_sizes;

int gh_space_sizes_get(const gh_space* space, gh_space_sizes* sizes);
/* leaves (Morton order), degrees (dim per leaf), mode_offsets (leaf_count+1),
 * mode_box (mode_count), term_offsets (mode_count+1), term_indices,
 * term_coefficients (term_count). Any pointer may be NULL to skip it. */
int gh_space_export(const gh_space* space, int64_t* leaves, int32_t* degrees,
                    int64_t* mode_offsets, int32_t* mode_box, int64_t* term_offsets,
                    int64_t* term_indices, double* term_coefficients);
/* Boundary faces (leaf index, axis, side) and interfaces (minus, plus, axis). */
int gh_space_faces(const gh_space* space, int64_t* boundary_leaf, int32_t* boundary_axis,
                   int32_t* boundary_side, int64_t* interface_minus, int64_t* interface_plus,
                   int32_t* interface_axis);

/* Borrowed view of a compiled space (arrays as returned by gh_space_export)
 * plus the reduced map of a plan on it: full -> reduced CSR and, for each full
 * dof, its reduced index or -1 when it is not a solve dof. */
typedef struct gh_space_view {
  int32_t dim;
  int64_t leaf_count;
  const int64_t* leaves;
  const int32_t* degrees;
  const int64_t* mode_offsets;
  const int32_t* mode_box;
  const int64_t* term_offsets;
  const int64_t* term_indices;
  const double* term_coefficients;
  int64_t full_dof_count;
  const int64_t* map_offsets;
  const int64_t* map_indices;
  const double* map_coefficients;
  const int64_t* reduced_index;
} gh_space_view;

typedef struct gh_transfer gh_transfer;

/* Reduced-space prolongation from `coarse` to `fine` (nested spaces of one
 * grid: same leaves with degrees >= the coarse ones, or the two children of a
 * coarse leaf). Entries are sorted by (fine, coarse); the V-cycle prolongs
 * with them and restricts with their transpose (src/multigrid.jl:614-857,
 * 1060-1087). */
int gh_transfer_create(const gh_grid* grid, const gh_space_view* coarse, const gh_space_view* fine,
                       gh_transfer** out);
void gh_transfer_destroy(gh_transfer* transfer);
int64_t gh_transfer_nnz(const gh_transfer* transfer);
int gh_transfer_export(const gh_transfer* transfer, int64_t* coarse_indices, int64_t* fine_indices,
                       double* coefficients);

/* Greedy coloring of blocks of dofs such that no item touches dofs of two
 * blocks of one color: block b holds dofs block_dofs[block_ptr[b] ..
 * block_ptr[b+1]), item i touches item_dofs[item_ptr[i] .. item_ptr[i+1])
 * (dofs in 0 .. dof_count - 1). Blocks are colored in order with the smallest
 * color no conflicting earlier block holds. Used to probe cell blocks and
 * operator diagonals with one apply per color and position (no reference
 * counterpart: the reference probes one unit vector per dof,
 * src/solve.jl:722-738). */
int gh_conflict_coloring(int64_t block_count, const int64_t* block_ptr, const int64_t* block_dofs,
                         int64_t item_count, const int64_t* item_ptr, const int64_t* item_dofs, int64_t dof_count,
                         int32_t* colors, int32_t* color_count);

#ifdef __cplusplus
}
#endif

#endif /* GRICO_HOST_H */
