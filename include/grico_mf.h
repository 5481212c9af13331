/*
 * grico_mf.h — device C ABI of the MI355X (gfx950) matrix-free hp-FEM backend.
 *
 * The entry points replace the CPU runtime of Grico.jl's compiled plans. Every
 * function cites the reference routine it stands in for (paths relative to the
 * Grico.jl source tree). A Julia binding would `ccall` these symbols (see
 * INTEGRATION.md); the Python package grico.jl_amd/gricoamd binds them with
 * ctypes.
 *
 * Conventions
 *  - Indices are 0-based (the Julia side subtracts 1 once, inside plan
 *    construction).
 *  - Descriptor arrays passed to gmf_plan_create are HOST pointers and are
 *    copied; the caller may free them afterwards.
 *  - Vector arguments of apply/solve entry points are DEVICE pointers
 *    (hipMalloc'd, caller-owned, FP64), on the device the plan was created on.
 *  - A workspace owns one HIP stream slot (default: the null stream, or the
 *    stream set with gmf_workspace_set_stream). Calls are synchronous with
 *    respect to the host unless stated otherwise. A plan may be shared by
 *    several workspaces; one workspace must not be used by two threads at once.
 *  - Every function returns a gmf_status; gmf_last_error() returns a message
 *    for the most recent failure on the calling thread.
 */
#ifndef GRICO_MF_H
#define GRICO_MF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gmf_status {
  GMF_OK = 0,
  GMF_EINVAL = -1,      /* wrong length / shape / argument (ArgumentError) */
  GMF_EALIAS = -2,      /* result aliases input (src/assembly.jl:450-453) */
  GMF_ENOCONV = -3,     /* Krylov method did not converge (src/solve.jl:430, 541) */
  GMF_EBREAKDOWN = -4,  /* CG non-positive direction (src/solve.jl:415-416) */
  GMF_EDIAGONAL = -5,   /* zero / non-finite Jacobi diagonal (src/solve.jl:740-751) */
  GMF_ENOTSPD = -6,     /* coarse matrix not SPD / singular (src/multigrid.jl:387-407) */
  GMF_EDEVICE = -7,     /* HIP runtime failure */
  GMF_EUNSUPPORTED = -8 /* shape outside what the kernels implement */
} gmf_status;

typedef struct gmf_plan gmf_plan;
typedef struct gmf_workspace gmf_workspace;

/* Per-point action forms lowered from cell accumulators on the host. */
enum {
  GMF_FORM_DIFFUSION = 0, /* a_grad = c_q * grad u, a_0 = 0: one coefficient c_q*W_q per point */
  GMF_FORM_GENERAL = 1    /* (a_0, a_grad) = M_q (u, grad u): (D+1)^2 coefficients per point */
};

/*
 * One batch of tensor-product cells sharing degrees, quadrature shape and
 * action form (the reference's kernel-signature batches, src/plans.jl:307-332).
 * Boxes and point sets are ordered axis 1 fastest (src/quadrature.jl:203-205).
 */
typedef struct gmf_bucket_desc {
  int32_t dim;       /* 1, 2 or 3 */
  int32_t modes[3];  /* N_a = p_a + 1 */
  int32_t points[3]; /* Q_a */
  int32_t form;      /* GMF_FORM_* */
  int64_t cell_count;
  /* Per axis a: N_a x Q_a row-major [mode][point] tables, axes concatenated:
   * values psi_i(xi_q) and reference-coordinate derivatives psi_i'(xi_q)
   * (src/integration.jl:2677-2742 before the J^-1 scaling). */
  const double* values;
  const double* derivatives;
  const double* inverse_jacobian; /* cell_count x dim (diagonal J^-1) */
  /* cell_count x Q x k, Q = prod Q_a: for DIFFUSION k = 1 (W_q c_q); for
   * GENERAL k = (D+1)^2 row-major [test channel][trial channel] of W_q M_q,
   * channel 0 = value, 1+a = physical derivative along axis a. */
  const double* point_data;
  /* cell_count x N, N = prod N_a (full box): >= 0 single global dof,
   * -1 inactive / constrained-to-zero mode, <= -2 multi-term entry -(v+2). */
  const int32_t* local_index;
  const double* local_coefficient; /* cell_count x N single-term coefficients, NULL = all 1 */
  int64_t multi_count;
  const int64_t* multi_offsets; /* multi_count + 1 */
  const int32_t* multi_indices;
  const double* multi_coefficients;
  /* Point-cloud batches (cut cells with their own quadrature, embedded surface
   * pieces; src/accumulators.jl:1022-1054 basis path, 1133-1214 one-sided
   * surface apply): when point_offsets is not NULL, item c owns points
   * point_offsets[c] .. point_offsets[c+1]-1 with biunit reference coordinates
   * ref_points (total x dim, row-major) and point_data rows (total x (D+1)^2,
   * GENERAL form, weights included); `points`, `values` and `derivatives` are
   * then ignored and the basis is evaluated at the points on the device. */
  const int64_t* point_offsets;
  const double* ref_points;
  /* Two-sided interface items (src/integration.jl:187-207 InterfaceValues,
   * src/accumulators.jl:1576-1616 interface_apply!): sides = 2 on a
   * point-cloud batch. Item c couples a minus leaf (modes[]) and a plus leaf
   * (plus_modes[]); its local_index row holds the minus box followed by the
   * plus box (N = prod modes + prod plus_modes), ref_points rows hold the
   * minus then the plus reference coordinates (total x 2 dim),
   * inverse_jacobian rows the minus then the plus diagonal (cell_count x
   * 2 dim), and point_data rows are (2(D+1))^2 row-major over the channels
   * s(D+1) + c (s = 0 minus, 1 plus; c = 0 value, 1 + a physical derivative).
   * sides = 0 or 1: one-sided items (plus_modes ignored). */
  int32_t sides;
  int32_t plus_modes[3];
  /* Optional 2x2x2 brick grouping of a tensor-product batch (NULL: consecutive
   * groups of eight cells, or none). Brick b holds the batch cells
   * brick_cells[8 b + c] (-1: no cell there) at position c = x + 2 y + 4 z
   * within the brick; each cell of the batch appears once. With it, batches
   * whose cells leave gaps in the 2x2x2 groups (cells cut by an embedded
   * boundary taken out, cells outside the domain dropped) still run on the
   * brick kernel. A traversal hint of the device plan with no counterpart in
   * the reference, whose cell traversal is src/plans.jl:368-424. */
  const int32_t* brick_cells;
  int64_t brick_count;
} gmf_bucket_desc;

/* Compiled plan: cell batches + algebraic constraint rows + reduced map
 * (src/plans.jl:113-145 AssemblyPlan; src/assembly.jl:19-27 _ReducedOperatorMap). */
typedef struct gmf_plan_desc {
  int32_t dim;
  int64_t full_dof_count;
  int32_t bucket_count;
  const gmf_bucket_desc* buckets;
  /* Constraint rows written after the cell pass (src/assembly.jl:1893-1920). */
  int64_t fixed_count;
  const int64_t* fixed_dofs; /* result[d] = x[d] */
  int64_t affine_count;      /* result[p] = x[p] - sum c_j x[j] */
  const int64_t* affine_pivots;
  const int64_t* affine_offsets;
  const int64_t* affine_indices;
  const double* affine_coefficients;
  int64_t mean_count; /* result[p] = sum c_j x[j] */
  const int64_t* mean_pivots;
  const int64_t* mean_offsets;
  const int64_t* mean_indices;
  const double* mean_coefficients;
  /* u_full = shift + P u_red, CSR over full dofs (src/assembly.jl:282-314). */
  int64_t reduced_dof_count;
  const int64_t* map_offsets; /* full_dof_count + 1 */
  const int64_t* map_indices;
  const double* map_coefficients;
  const double* shift;      /* full_dof_count */
} gmf_plan_desc;

typedef struct gmf_solve_info {
  int64_t iterations;
  int64_t operator_applications;
  int64_t preconditioner_applications;
  double initial_residual_norm;
  double final_residual_norm;
  double rhs_norm;
  int32_t converged;
} gmf_solve_info;

/* Opaque preconditioner: z = M^-1 r on reduced vectors (src/solve.jl:32
 * AbstractPreconditioner, 641-670 _compile_preconditioner, 753-774
 * _apply_preconditioner!). NULL means the identity. */
typedef struct gmf_preconditioner gmf_preconditioner;

const char* gmf_last_error(void);

/* Device properties of the current HIP device (name buffer >= 64 bytes). */
gmf_status gmf_device_info(char* name, int32_t* compute_units, int64_t* global_memory_bytes);

/* Plan lifetime: replaces compile(problem)'s frozen AssemblyPlan
 * (src/plans.jl:170-231); tables are uploaded to HBM once. */
gmf_status gmf_plan_create(const gmf_plan_desc* desc, gmf_plan** plan);
gmf_status gmf_plan_destroy(gmf_plan* plan);
gmf_status gmf_plan_sizes(const gmf_plan* plan, int64_t* full_dof_count, int64_t* reduced_dof_count,
                          int64_t* device_bytes);

/* Human-readable summary of the device plan (batches, kernel paths, brick
 * colors, bytes), NUL-terminated into buf (diagnostics; no reference
 * counterpart). */
gmf_status gmf_plan_describe(const gmf_plan* plan, char* buf, int64_t capacity);

/* Kernel family and launch shape one bucket runs with, picked at plan
 * creation by timing the candidates (no reference counterpart: the reference
 * runs every batch through one CPU loop). kernel: 0 empty bucket, 1 brick
 * kernel, 2 per-cell tensor kernel (param0 = lanes per workgroup), 3 generic
 * tensor kernel (param0 = lanes per cell), 4 point-cloud kernel (param0 =
 * points per work unit, param1 = LDS budget in KB), 5 face-tensor kernel,
 * 6 point-cloud items through local matrices built from the point data at
 * plan creation (param0 = local modes per item).
 * candidates = shapes timed (0: defaults kept, GMF_TUNE=0); best_us /
 * default_us = time per launch of the chosen / the default shape. */
typedef struct gmf_bucket_tuning {
  int32_t kernel;
  int32_t param0;
  int32_t param1;
  int32_t candidates;
  double best_us;
  double default_us;
} gmf_bucket_tuning;
gmf_status gmf_plan_tuning(const gmf_plan* plan, int32_t bucket, gmf_bucket_tuning* out);

/* Workspace: per-caller scratch tied to one plan (src/assembly.jl:388-441,
 * 566-584 OperatorWorkspace / _ReducedOperatorWorkspace). `stream` is a
 * hipStream_t (NULL = null stream). */
gmf_status gmf_workspace_create(gmf_plan* plan, gmf_workspace** ws);
gmf_status gmf_workspace_destroy(gmf_workspace* ws);
gmf_status gmf_workspace_set_stream(gmf_workspace* ws, void* stream);

/* y = A x on the full layout with constraint rows (src/assembly.jl:540-564 apply!). */
gmf_status gmf_apply(gmf_workspace* ws, const double* x, double* y);
/* Same, enqueued on the workspace stream without a host synchronization. */
gmf_status gmf_apply_async(gmf_workspace* ws, const double* x, double* y);
/* y = A x on the full layout without constraint rows (src/assembly.jl:603-622 _apply_physical!). */
gmf_status gmf_apply_physical(gmf_workspace* ws, const double* x, double* y);
/* y_red = P^T A P x_red (src/assembly.jl:905-915 _reduced_apply!). */
gmf_status gmf_reduced_apply(gmf_workspace* ws, const double* x_red, double* y_red);

/* y_red = P^T A P x_red like gmf_reduced_apply, allocating nothing. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops). The workspace scratch
 * comes from a fixed 32 MiB device array that the library holds per device
 * until the process ends; plans that need more (brick halo slots, plus twice
 * the full dof count when the reduced map is not a selection) or have more
 * than 64 cell batches are rejected with GMF_EUNSUPPORTED. The plan's device
 * must be current. Asynchronous on `stream` (NULL: the null stream); calls that
 * overlap on one device share the array and race. */
gmf_status gmf_reduced_apply_noalloc(const gmf_plan* plan, const double* x, double* y, void* stream);

/* The same operator on FP32 vectors, the level operator of the FP32 V-cycle:
 * FP32 cell kernels where every bucket of the plan has them (3D tensor cells
 * with N = Q <= 8 and the diffusion form, selection reduced maps), otherwise
 * the FP64 action of the widened input rounded to FP32. No reference
 * counterpart (the reference computes in the vector element type,
 * src/assembly.jl:905). The first call on a workspace allocates its FP32
 * copies of the per-cell arrays. */
gmf_status gmf_reduced_apply_f32(gmf_workspace* ws, const float* x_red, float* y_red);
/* Physical right-hand side from lowered point loads (src/assembly.jl:488-507
 * rhs!, 586-601 _rhs_physical!): y = sum over cells of the projections
 * sum_q [phi_i(x_q) v_q + grad phi_i(x_q) . g_q]. `loads` holds, for the plan's
 * tensor-product cells in bucket order and each cell's points, the channels
 * (W v) or (W v, W g_1, .., W g_D) (channels = 1 or 1 + D), W the quadrature
 * weight: the load callbacks evaluated on the host, as the operator action is
 * lowered to point tables. Point-cloud batches are skipped (their loads:
 * gmf_rhs_faces). Device pointers; synchronous. */
gmf_status gmf_rhs_physical(gmf_workspace* ws, const double* loads, int32_t channels, double* y);
/* Point-cloud loads added into y (full layout): cut-cell loads
 * (src/accumulators.jl:1022-1054 basis path of cell_rhs), embedded-surface and
 * boundary-face loads (1230-1274 _accumulate_one_sided_rhs!) and interface
 * loads (1666-1698 interface_rhs!). For the plan's point-cloud batches in
 * bucket order, per item, point and side (minus then plus for interfaces) the
 * channels (W v, W g_1, .., W g_D) of the load callbacks evaluated on the
 * host; batches on the sum-factorized face path project them with the face
 * kernel, the others with the point-cloud kernel. Device pointers;
 * synchronous. */
gmf_status gmf_rhs_faces(gmf_workspace* ws, const double* loads, double* y);
/* State channels at the points of the plan's tensor-product cells (bucket
 * order, cells x Q x channels): the value of the full-layout u and, for
 * channels = 1 + D, its physical gradient - the StateChannels handed to
 * cell_residual_accumulate / cell_tangent_accumulate
 * (src/accumulators.jl:752-787 _tensor_state_channels, 2015-2148 residual and
 * tangent tensor kernels). With gmf_rhs_physical (residual loads) and
 * gmf_plan_set_point_data (tangent re-lowered at the state) it carries the
 * matrix-free residual! / tangent_apply! (src/assembly.jl:997-1080).
 * Point-cloud items are not covered (GMF_EUNSUPPORTED). Device pointers. */
gmf_status gmf_interpolate_channels(gmf_workspace* ws, const double* u_full, int32_t channels, double* out);
/* Replace the lowered point action of tensor bucket `bucket` (host array of
 * the plan descriptor's point_data shape and form): the tangent of a residual
 * problem lowered at a new state (src/assembly.jl:1050-1080 tangent_apply!).
 * Not concurrent with calls that use the plan. */
gmf_status gmf_plan_set_point_data(gmf_plan* plan, int32_t bucket, const double* point_data);
/* b_red = P^T (rhs_full - A_phys shift) (src/assembly.jl:917-933 _reduced_rhs!). */
gmf_status gmf_reduced_rhs(gmf_workspace* ws, const double* rhs_full, double* b_red);
/* Verification sums (src/verification.jl:201-230 _l2_error_components):
 * sums[0] = sum_q w_q (u_h(x_q) - e_q)^2, sums[1] = sum_q w_q e_q^2, u_h from the
 * full-layout coefficients u_full. Weights and exact values per point: the
 * tensor-product cells' points in the gmf_rhs_physical `loads` order, then, in
 * bucket order, the points of every one-sided point-cloud batch (cut cells;
 * zero weights leave out surface pieces and boundary faces, whose points are
 * skipped on the face path). Two-sided batches take no values. sums is a host
 * array of two doubles. */
gmf_status gmf_l2_error(gmf_workspace* ws, const double* u_full, const double* weights, const double* exact,
                        double* sums);
/* x_full = [shift] + P x_red (src/assembly.jl:851-874 _expand_reduced!). */
gmf_status gmf_expand_reduced(gmf_workspace* ws, const double* x_red, double* x_full, int32_t include_shift);
/* y_red = P^T y_full (src/assembly.jl:876-891 _project_reduced!). */
gmf_status gmf_project_reduced(gmf_workspace* ws, const double* y_full, double* y_red);
/* Diagonal of the physical operator on the full layout, scattered with c^2
 * (src/assembly.jl:624-639 _diagonal_physical!, src/accumulators.jl:1056-1083). */
gmf_status gmf_diagonal_physical(gmf_workspace* ws, double* diag_full);
/* Exact diagonal of P^T A P (src/assembly.jl:662-675 _reduced_diagonal! and the
 * probing fallback src/solve.jl:722-738, computed without probing). */
gmf_status gmf_reduced_diagonal(gmf_workspace* ws, double* diag_red);

/* Jacobi preconditioner from an inverse diagonal (device vector of length
 * reduced_dof_count, copied) (src/solve.jl:651-656, 761-774). */
gmf_status gmf_precond_create_jacobi(gmf_workspace* ws, const double* inverse_diagonal,
                                     gmf_preconditioner** out);
/* Jacobi preconditioner applied in FP32: the inverse diagonal (device FP64,
 * reduced_dof_count) is rounded to FP32 and z = D^-1 r is formed in FP32, then
 * widened (the single-precision preconditioner of an FP64 Krylov method; the
 * reference's Jacobi, src/solve.jl:651-656, 761-774, in FP64 only). */
gmf_status gmf_precond_create_jacobi_f32(gmf_workspace* ws, const double* inverse_diagonal,
                                         gmf_preconditioner** out);
gmf_status gmf_precond_destroy(gmf_preconditioner* precond);
/* z = M^-1 r on the workspace stream. */
gmf_status gmf_precond_apply(gmf_preconditioner* precond, const double* r, double* z);
/* Z = M^-1 R for k columns of the preconditioner's reduced size (column-major
 * device blocks), what the block solves apply: FP64 multigrid V-cycles with
 * Jacobi or Chebyshev smoothing run all columns per launch (a dense coarse
 * level as one matrix product), other preconditioners column by column
 * (src/solve.jl:753-774 _apply_preconditioner! per right-hand side). */
gmf_status gmf_precond_apply_block(gmf_preconditioner* precond, const double* R, double* Z, int32_t k);

/* Preconditioned CG on the reduced system from x (in: initial guess if
 * use_initial, else zero), stopping at ||r|| <= max(atol, rtol ||b||)
 * (src/solve.jl:376-431 _cg_solve). */
gmf_status gmf_cg_solve(gmf_workspace* ws, gmf_preconditioner* precond, const double* b, double* x,
                        int32_t use_initial, double rtol, double atol, int64_t maxiter,
                        gmf_solve_info* info);
/* Several right-hand sides at once (no single reference counterpart: the
 * batched form of k calls of src/solve.jl:376-431 _cg_solve on one operator).
 * Blocks are reduced_dof_count x k, column j at offset j * reduced_dof_count,
 * 1 <= k <= 4096. gmf_reduced_apply_block: Y_j = P^T A P X_j (src/assembly.jl:
 * 905-915 per column; point-cloud batches take every column in one launch).
 * gmf_cg_solve_block: CG on every column in lockstep with its own scalars, the
 * iteration of gmf_cg_solve per column; a converged column is frozen at the
 * iterate where it converged; info holds k entries. GMF_ENOCONV when a column
 * has not converged after maxiter iterations (info still filled). */
gmf_status gmf_reduced_apply_block(gmf_workspace* ws, const double* X, double* Y, int32_t k);
gmf_status gmf_cg_solve_block(gmf_workspace* ws, gmf_preconditioner* precond, const double* B, double* X,
                              int32_t k, int32_t use_initial, double rtol, double atol, int64_t maxiter,
                              gmf_solve_info* info);
/* Restarted right-preconditioned FGMRES with modified Gram-Schmidt and Givens
 * rotations (src/solve.jl:433-614 _fgmres_solve). */
gmf_status gmf_fgmres_solve(gmf_workspace* ws, gmf_preconditioner* precond, const double* b, double* x,
                            int32_t use_initial, int32_t restart, double rtol, double atol,
                            int64_t maxiter, gmf_solve_info* info);

/* ---- hp geometric multigrid (src/multigrid.jl) --------------------------- */

enum {
  GMF_SMOOTHER_JACOBI = 0,   /* damped Jacobi sweeps, the reference's SPD smoother (src/multigrid.jl:943-956) */
  GMF_SMOOTHER_CHEBYSHEV = 1, /* Chebyshev polynomial in D^-1 A on [lambda_max / range, lambda_max]; no reference counterpart */
  GMF_SMOOTHER_GMRES = 2      /* one fixed cycle of min(steps, restart) Jacobi right-preconditioned Arnoldi steps,
                                 damped update: the reference's smoother for general operators
                                 (src/multigrid.jl:963-1035); FP64 levels only */
};

/* Reduced-space prolongation from level i to level i+1 as COO entries sorted
 * by (fine, coarse) (src/multigrid.jl:115-121 _ReducedTransfer); HOST arrays,
 * copied. Restriction is the transpose (src/multigrid.jl:1060-1087). */
typedef struct gmf_transfer_desc {
  int64_t nnz;
  const int64_t* coarse_indices;
  const int64_t* fine_indices;
  const double* coefficients;
} gmf_transfer_desc;

/* Cell blocks of one multigrid level for block smoothing (no reference
 * counterpart; the reference smooths with the point diagonal only): block b
 * holds the reduced dofs dofs[offsets[b] .. offsets[b+1]) (distinct, at most
 * 512) and `inverse` the row-major inverses of the blocks B_b = R_b A R_b^T
 * one after another (m_b^2 values each; symmetric). The level's smoother then
 * preconditions with S = D_u^-1 + sum_b R_b^T B_b^-1 R_b, D_u^-1 the Jacobi
 * inverse diagonal on the dofs no block covers. HOST arrays, copied. */
typedef struct gmf_block_desc {
  int64_t count;
  const int64_t* offsets;
  const int32_t* dofs;
  const double* inverse;
} gmf_block_desc;

/* The blocks B_b = R_b A R_b^T of a workspace's reduced operator (row-major,
 * m_b^2 values per block one after another into `out`, a HOST array) by
 * probing: for every color and every local position j, one reduced apply of
 * the sum of the unit vectors of the j-th dofs of that color's blocks. Blocks
 * of one color must be independent: no cell of the plan may touch dofs of two
 * of them (otherwise the probes mix; the result is then wrong, not an error).
 * colors[b] in 0 .. color_count - 1. `blocks.inverse` is not read. */
gmf_status gmf_reduced_blocks(gmf_workspace* ws, const gmf_block_desc* blocks, const int32_t* colors,
                              int32_t color_count, double* out);

/* V-cycle hierarchy (src/multigrid.jl:51-96 GeometricMultigridPreconditioner,
 * 309-345 _compile_geometric_multigrid). levels[0] is the coarsest level and
 * levels[level_count-1] the workspace of the operator being preconditioned; the
 * preconditioner belongs to that workspace. Each level's plan is the problem
 * rediscretized on that level's space. */
typedef struct gmf_mg_desc {
  int32_t level_count; /* >= 2 */
  gmf_workspace* const* levels;
  const gmf_transfer_desc* transfers; /* level_count - 1; transfers[i]: level i -> i+1 */
  int32_t smoother;                   /* GMF_SMOOTHER_* */
  int32_t pre_steps;                  /* Jacobi sweeps / Chebyshev applications before restriction */
  int32_t post_steps;
  double damping;                     /* Jacobi omega (reference default 0.8) */
  int32_t chebyshev_degree;           /* operator applications per Chebyshev application */
  double chebyshev_range;             /* lambda_min = lambda_max / range */
  int32_t eigen_iterations;           /* Lanczos steps estimating lambda_max(D^-1 A) per level */
  int64_t coarse_direct_limit;        /* dense SPD solve when n_coarse <= limit (reference default 512) */
  double coarse_rtol;                 /* else Jacobi-CG to these tolerances (src/multigrid.jl:891-924) */
  double coarse_atol;
  int64_t coarse_maxiter;
  int32_t graph; /* 1: replay the V-cycle as a captured HIP graph (dense coarse solve only) */
  /* Level arithmetic: 64 (or 0) FP64 throughout; 32: smoothers, transfers and
   * the level operators in FP32 (FP32 cell kernels where the level's buckets
   * have them, else the FP64 apply of the widened vector, rounded), while the
   * coarse solve, the Chebyshev eigenvalue estimates and the Jacobi diagonals
   * are computed in FP64. The outer Krylov method stays FP64: z = M^-1 r
   * rounds r to FP32 and widens the FP32 V-cycle result. No reference
   * counterpart (the reference V-cycle runs in the vector element type). */
  int32_t precision;
  int32_t smoother_restart; /* GMRES smoother Krylov length (reference default 8); >= max(pre, post) */
  /* 1: SPD operator (Cholesky dense coarse inverse, CG coarse solve); 0: general
   * (LU dense coarse inverse, FGMRES coarse solve with restart min(30, n))
   * (src/multigrid.jl:382-408, 891-924). */
  int32_t spd;
  /* Optional HOST arrays, one per level (NULL, or NULL entries: computed on
   * the device): the reduced diagonal of each level's operator. Levels whose
   * reduced map is not a selection (affine Dirichlet relations, mean rows)
   * need it, since the device's exact diagonal covers selection maps only;
   * the reference probes such diagonals (src/solve.jl:722-738). */
  const double* const* level_diagonals;
  /* Optional, level_count entries (NULL, or entries with count 0: point
   * smoothing): cell blocks of each level's block smoother (see
   * gmf_block_desc; Jacobi and Chebyshev smoothers; the coarsest level's entry
   * is not used). The Chebyshev bounds are then estimated for S A. */
  const gmf_block_desc* level_blocks;
} gmf_mg_desc;

/* Compile the V-cycle: level Jacobi diagonals, Chebyshev bounds, transfer CSR
 * tables and the coarse solver (dense Cholesky-based inverse of the probed
 * coarse matrix, rejected with GMF_ENOTSPD if not symmetric positive definite,
 * src/multigrid.jl:382-408). z = M^-1 r runs one V-cycle from z = 0
 * (src/multigrid.jl:858-889). */
gmf_status gmf_precond_create_mg(const gmf_mg_desc* desc, gmf_preconditioner** out);

/* Estimated largest eigenvalue of D^-1 A used by level `level`'s Chebyshev
 * smoother (0 for the coarse level or the Jacobi smoother). Diagnostics. */
gmf_status gmf_mg_level_info(const gmf_preconditioner* mg, int32_t level, int64_t* reduced_dofs, double* lambda_max);

/* ---- learned per-cell block correction ---------------------------------------
 * No reference counterpart: a small per-cell MLP (BF16 weights on MFMA, FP32
 * accumulation) used as an additional correction inside FGMRES, which accepts
 * variable preconditioners (src/solve.jl:80-88), through the preconditioner
 * seam src/solve.jl:32, 641-670, 753-774. */
typedef struct gmf_mlp gmf_mlp;

typedef struct gmf_mlp_desc {
  int32_t input, hidden1, hidden2, output; /* layer widths */
  /* Initial FP32 parameters, packed [W1 (input x hidden1, row-major), b1,
   * W2 (hidden1 x hidden2), b2, W3 (hidden2 x output), b3]; y = W^T x + b per layer,
   * ReLU between layers. HOST array, copied. */
  const float* parameters;
} gmf_mlp_desc;

gmf_status gmf_mlp_create(const gmf_mlp_desc* desc, gmf_mlp** out);
gmf_status gmf_mlp_destroy(gmf_mlp* mlp);
gmf_status gmf_mlp_set_stream(gmf_mlp* mlp, void* stream);
gmf_status gmf_mlp_parameter_count(const gmf_mlp* mlp, int64_t* count);
/* Current FP32 master parameters into a HOST array (packed as in gmf_mlp_desc). */
gmf_status gmf_mlp_get_parameters(gmf_mlp* mlp, float* parameters);
/* y = net(x) for `rows` FP32 rows (DEVICE arrays). */
gmf_status gmf_mlp_forward(gmf_mlp* mlp, const float* x, float* y, int64_t rows);
/* Mean squared error of net(x) against t and its parameter gradient (HOST,
 * packed) for one batch (DEVICE x, t). */
gmf_status gmf_mlp_gradient(gmf_mlp* mlp, const float* x, const float* t, int64_t rows, double* loss, float* gradient);
/* Adam on consecutive batches of `batch` rows (cycling through
 * floor(samples / batch) batches) for `steps` steps; loss_history (HOST,
 * steps entries) may be NULL. */
gmf_status gmf_mlp_train(gmf_mlp* mlp, const float* x, const float* t, int64_t samples, int64_t batch, int64_t steps,
                         float learning_rate, float beta1, float beta2, float eps, double* loss_history);
/* Cell patches of reduced vectors for a single-batch plan: x[p][i] =
 * r[dof(cell_p, i)] / ||r|_cell_p||, scale[p] = ||r|_cell_p||, and with e,
 * t[p][i] = e[dof(cell_p, i)] * cell_weight[cell_p] / scale[p]. `cells`
 * (DEVICE, count entries) may be NULL for all cells in order. */
gmf_status gmf_cell_patches(gmf_workspace* ws, const double* r, const double* e, const int64_t* cells, int64_t count,
                            const float* cell_weight, float* x, float* t, float* scale);
/* z = M(r) = z0 + P_cells(net(patches(r - A z0)) * scale / cell_weight), z0 =
 * inner(r) (or 0 when inner is NULL), the per-cell outputs averaged on shared
 * dofs. cell_weight: HOST array of one positive weight per cell (copied). */
gmf_status gmf_precond_create_learned(gmf_workspace* ws, gmf_preconditioner* inner, gmf_mlp* net,
                                      const float* cell_weight, gmf_preconditioner** out);

/* Training pairs for the learned preconditioner from GMG-CG solves with its
 * inner preconditioner V (required): per solve, a seeded reference solution
 * x* ~ N(0, 1) and b = A x*; at each CG iterate x_k (at most `iterations` per
 * solve), y = x_k + V(b - A x_k), residual r' = b - A y (row of `residuals`,
 * DEVICE, pairs x n_red, FP32) and energy e'^T r' with e' = x* - y (HOST,
 * pairs). Runs on the preconditioner's workspace stream. */
gmf_status gmf_learned_collect(gmf_preconditioner* learned, int64_t pairs, int32_t iterations, uint64_t seed,
                               float* residuals, double* energies);

/* Energy-norm training of the learned preconditioner's network: for each
 * residual r'_k (DEVICE, pairs x n_red FP32, row k = pair k) left by the inner
 * preconditioner, minimise ||e'_k - c(r'_k)||_A^2 = e'^T A e' - 2 c^T r' + c^T A c
 * with one Adam step per residual (cycling). energies (HOST) = e'_k^T r'_k;
 * history (HOST, steps entries, may be NULL) records ||e' - c||_A^2 / ||e'||_A^2. */
gmf_status gmf_learned_train(gmf_preconditioner* learned, const float* residuals, const double* energies,
                             int64_t pairs, int64_t steps, float learning_rate, float beta1, float beta2, float eps,
                             double* history);

/* The same objective on one residual without a parameter update: loss =
 * ||e' - c(r')||_A^2 / energy (energy = e'^T r' > 0) and its gradient over the
 * packed FP32 parameters (HOST array of gmf_mlp_parameter_count entries; may be
 * NULL). residual: DEVICE, n_red. The gradient checks of the training path use it. */
gmf_status gmf_learned_gradient(gmf_preconditioner* learned, const double* residual, double energy, double* loss,
                                float* grad_host);

/* Timed repetitions of gmf_apply on the workspace stream with HIP events
 * around each stage: average milliseconds of the cell-kernel pass and of the
 * whole apply (the measurement role of src/diagnostics.jl:21-59
 * operator_diagnostics). */
gmf_status gmf_apply_timed(gmf_workspace* ws, const double* x, double* y, int32_t repetitions,
                           double* cell_ms, double* total_ms);

/* dst = src for n doubles with the library's plain streaming copy kernel on
 * the workspace stream (vector utility; also the known-size kernel used to
 * calibrate HBM byte counters). No reference counterpart. */
gmf_status gmf_vector_copy(gmf_workspace* ws, double* dst, const double* src, int64_t n);

/* y = a x + b y and *out = x . y on reduced-length device vectors (vector
 * utilities for callers that drive their own iterations, e.g. training-data
 * generation for the learned correction). No reference counterpart. */
gmf_status gmf_vector_axpby(gmf_workspace* ws, double* y, double a, const double* x, double b, int64_t n);
gmf_status gmf_vector_dot(gmf_workspace* ws, const double* x, const double* y, int64_t n, double* out);

/* Jacobi inverse diagonal with validation: rejects zero / non-finite entries
 * (src/solve.jl:740-751). */
gmf_status gmf_invert_diagonal(gmf_workspace* ws, const double* diag, double* inverse_diagonal, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* GRICO_MF_H */
