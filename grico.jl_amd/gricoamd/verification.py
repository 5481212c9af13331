"""L2 errors against an exact solution on the compiled cells.

Reference: src/verification.jl:201-230 (_l2_error_components: sqrt of
sum_q W_q (u_h(x_q) - u(x_q))^2 over the plan's cells), l2_error 270,
relative_l2_error 297. Evaluated on the host from the solved state with the
plan's own point sets and sum-factorized tables.
"""

import numpy as np


def _field_values(plan, coefficients, bucket):
    """u_h at every point of every cell of one bucket."""
    box = np.zeros((bucket.cell_count, bucket.n_modes))
    contrib = bucket.term_coef * coefficients[bucket.term_dof]
    np.add.at(box, (bucket.term_cell, bucket.term_box), contrib)
    return bucket.interpolate(box)


def l2_error(state, exact):
    plan = state.plan
    total = 0.0
    for b in list(plan.buckets) + list(getattr(plan, "cloud_buckets", [])):
        uh = _field_values(plan, state.coefficients, b)
        ue = np.broadcast_to(np.asarray(exact(tuple(b.phys_points)), dtype=np.float64), uh.shape)
        total += float(np.sum(b.weights * (uh - ue) ** 2))
    return float(np.sqrt(total))


def relative_l2_error(state, exact):
    plan = state.plan
    num = 0.0
    den = 0.0
    for b in list(plan.buckets) + list(getattr(plan, "cloud_buckets", [])):
        uh = _field_values(plan, state.coefficients, b)
        ue = np.broadcast_to(np.asarray(exact(tuple(b.phys_points)), dtype=np.float64), uh.shape)
        num += float(np.sum(b.weights * (uh - ue) ** 2))
        den += float(np.sum(b.weights * ue ** 2))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def device_l2_error(state, exact):
    """l2_error with the interpolation and the weighted sum on the device
    (tensor and cut cells)."""
    from . import ops
    err, _ = ops.l2_sums(state.plan, state.coefficients, exact)
    return float(np.sqrt(err))


def device_relative_l2_error(state, exact):
    from . import ops
    err, ref = ops.l2_sums(state.plan, state.coefficients, exact)
    return float(np.sqrt(err / ref)) if ref > 0 else float(np.sqrt(err))
