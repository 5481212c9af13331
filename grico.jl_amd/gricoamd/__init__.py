"""MI355X matrix-free hp-FEM backend: host-side mirror of Grico.jl's API.

The names follow the reference's public surface (src/Grico.jl exports):
Domain, HpSpace, SpaceOptions, FullTensorBasis, TrunkBasis, UniformDegree,
AxisDegrees, ByLeafDegrees, DegreePlusQuadrature, ScalarField, AffineProblem,
SPD, TestChannels, TraceTestChannels, add_cell_accumulator,
add_boundary_accumulator, add_interface_accumulator, jump, average, add_constraint, Dirichlet, MeanValue,
BoundaryFace, LOWER, UPPER, compile, apply, solve, CGSolver, FGMRESSolver,
JacobiPreconditioner, GeometricMultigridPreconditioner, l2_error,
relative_l2_error, AdaptivityPlan, transition, transfer_state. Julia's `!`
mutators map to `*_into` functions (request_h_refinement! -> request_h_refinement).
"""

from .mesh import Domain, uniform_refine
from .space import (AxisDegrees, ByLeafDegrees, DegreePlusQuadrature, FullTensorBasis, HpSpace, SpaceOptions,
                    StoredDegrees, TrunkBasis, UniformDegree, scalar_dof_count)
from .problem import (LOWER, UPPER, AffineProblem, BoundaryFace, Dirichlet, MeanValue, NonsymmetricOperator, ScalarField, SPD,
                      TestChannels, TraceTestChannels, TraceTrialChannels, TracePair, TrialChannels,
                      add_boundary_accumulator, add_cell_accumulator, add_constraint, add_interface_accumulator,
                      average, cell_size, component, gradient, inner, jump, minus, normal, normal_component, plus,
                      value, weight)
from .plan import AssemblyPlan, compile
from .ops import (apply, apply_into, apply_physical, diagonal_physical, expand_reduced, project_reduced,
                  reduced_apply, reduced_diagonal)
from .solve import (AutoLinearSolver, CGSolver, ConvergenceError, FGMRESSolver, JacobiPreconditioner, State, solve,
                    solve_block, reduced_apply_block,
                    solve_reduced)
from .verification import device_l2_error, device_relative_l2_error, l2_error, relative_l2_error
from .multigrid import GeometricMultigridPreconditioner, MultigridHierarchy, compile_geometric_multigrid
from .embedded import FiniteCellExtension, ImplicitRegion, PhysicalDomain, PhysicalMeasure, SampledSurface
from .adaptivity import (AdaptivityLimits, AdaptivityPlan, SpaceTransition, evaluate, interface_jump_indicators,
                         coefficient_coarsening_indicators, transfer_state,
                         transition)

__all__ = [n for n in dir() if not n.startswith("_")]
