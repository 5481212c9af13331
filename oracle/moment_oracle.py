"""ORACLE (test infrastructure only): numpy restatement of the reference's
non-negative moment fitting of cut-cell quadratures (src/regions.jl:519-708),
step for step and with its tolerances:

- moment matrix: products of Legendre values P_0..P_q per axis at every
  candidate point, for every mode of the full tensor box 0..q_a
  (_finite_cell_moment_matrix 538-561, _legendre_values! polynomials.jl:294-307);
- NNLS: the Lawson-Hanson active-set loop of _solve_nonnegative_least_squares
  (641-694), each passive least-squares problem solved for the minimum-norm
  solution through the SVD with singular values below tol * s_max dropped
  (_least_squares_min_norm 709-725);
- point elimination (_point_elimination 563-619, _removable_weight_indices
  621-630): keep the positive weights, cut to the `rows` largest, then drop
  small weights one point at a time while the moments stay fitted, down to
  prod(q) points.

Only tests/ use this module. It is pinned there by what the mathematics
fixes: the NNLS agrees with scipy's Lawson-Hanson on problems with a unique
solution, fitted rules reproduce the candidate moments to the tolerance with
non-negative weights on a subset of the candidate points, and on tiny inputs
brute force over all point subsets confirms that the elimination stops only
where no single removal keeps an exact fit.
"""

import numpy as np

MOMENT_TOLERANCE_SCALE = 1.0e3      # src/regions.jl:519
WEIGHT_PRUNE_RATIO = 1.0e-8         # src/regions.jl:520
MAX_ELIMINATION_ITERATIONS = 1024   # src/regions.jl:521


def tolerance():
    return MOMENT_TOLERANCE_SCALE * np.finfo(np.float64).eps


def legendre_values(x, degree):
    """P_0..P_degree at x by the three-term recurrence (polynomials.jl:294-307)."""
    v = np.zeros(degree + 1)
    v[0] = 1.0
    if degree >= 1:
        v[1] = x
    for n in range(2, degree + 1):
        inv_n = 1.0 / n
        v[n] = (2 * n - 1) * inv_n * x * v[n - 1] - (n - 1) * inv_n * v[n - 2]
    return v


def box_modes(degrees):
    """Modes of the full tensor box 0..q_a, first axis fastest."""
    grids = np.meshgrid(*[np.arange(q + 1) for q in degrees][::-1], indexing="ij")
    return np.stack([g.reshape(-1) for g in grids[::-1]], axis=1)


def moment_matrix(points, degrees):
    """modes x points matrix of Legendre products (src/regions.jl:538-561)."""
    modes = box_modes(degrees)
    M = np.empty((len(modes), len(points)))
    for j, p in enumerate(points):
        vals = [legendre_values(p[a], degrees[a]) for a in range(len(degrees))]
        for i, m in enumerate(modes):
            v = 1.0
            for a in range(len(degrees)):
                v *= vals[a][m[a]]
            M[i, j] = v
    return M


def least_squares_min_norm(A, b, tol):
    """Minimum-norm least squares through the SVD, singular values at most
    tol * s_max dropped (src/regions.jl:709-725)."""
    if A.size == 0:
        return np.zeros(A.shape[1])
    U, S, Vt = np.linalg.svd(A, full_matrices=False)
    if S.size == 0:
        return np.zeros(A.shape[1])
    cutoff = tol * S[0]
    rb = U.T @ b
    scaled = np.where(S > cutoff, rb / np.where(S > cutoff, S, 1.0), 0.0)
    return Vt.T @ scaled


def nnls(A, b, tol):
    """(solution, relative residual) of min |A w - b|, w >= 0, by the
    reference's active-set loop (src/regions.jl:641-707)."""
    n = A.shape[1]
    w = np.zeros(n)
    passive = np.zeros(n, dtype=bool)
    gradient = A.T @ b
    max_it = max(8 * n, 32)
    it = 0
    while it < max_it:
        free = np.nonzero(~passive)[0]
        cand = -1
        best = 0.0
        for i in free:  # first index of the largest positive gradient
            if gradient[i] > best:
                best = gradient[i]
                cand = i
        if cand < 0 or not gradient[cand] > tol:
            break
        passive[cand] = True
        while True:
            idx = np.nonzero(passive)[0]
            sol = least_squares_min_norm(A[:, idx], b, tol)
            trial = np.zeros(n)
            trial[idx] = sol
            if np.all(sol > tol):
                w = trial
                break
            alpha = 1.0
            for i in idx:
                if trial[i] > tol:
                    continue
                den = w[i] - trial[i]
                if abs(den) <= tol:
                    continue
                alpha = min(alpha, w[i] / den)
            w = w + alpha * (trial - w)
            for i in idx:
                if w[i] <= tol:
                    w[i] = 0.0
                    passive[i] = False
            if not np.any(passive):
                break
        gradient = A.T @ (b - A @ w)
        it += 1
    res = b - A @ w
    return w, float(np.linalg.norm(res) / max(np.linalg.norm(b), 1.0))


def removable_weight_indices(weights, min_points, tol):
    """src/regions.jl:621-630."""
    if len(weights) <= min_points:
        return []
    threshold = max(tol, WEIGHT_PRUNE_RATIO * np.max(weights))
    rem = [i for i, w in enumerate(weights) if w <= threshold]
    if not rem:
        return [int(np.argmin(weights))]
    return sorted(rem, key=lambda i: weights[i])


def point_elimination(A, moments, min_points, tol):
    """Indices into the candidate points and their weights, or None when the
    candidate rule is kept (src/regions.jl:563-619)."""
    full_w, full_res = nnls(A, moments, tol)
    if not full_res <= tol:
        return None
    current = [i for i in range(A.shape[1]) if full_w[i] > tol]
    if not current:
        return None
    weights = full_w[current]
    if len(current) > A.shape[0]:
        order = np.argsort(-weights, kind="stable")
        current = [current[i] for i in order[:A.shape[0]]]
        weights, res = nnls(A[:, current], moments, tol)
        if not res <= tol:
            return None
    prev, prev_w = list(current), np.array(weights)
    for _ in range(MAX_ELIMINATION_ITERATIONS):
        if len(current) <= min_points:
            break
        rem = removable_weight_indices(weights, min_points, tol)
        if not rem:
            break
        accepted = False
        for li in rem:
            trial = [current[i] for i in range(len(current)) if i != li]
            if len(trial) < min_points:
                continue
            tw, res = nnls(A[:, trial], moments, tol)
            if not res <= tol:
                continue
            current, weights = trial, tw
            prev, prev_w = list(current), np.array(weights)
            accepted = True
            break
        if not accepted:
            break
    return prev, prev_w


def moment_fit(points, weights, shape):
    """(points, weights) of the fitted rule of a candidate rule on the biunit
    cell (src/regions.jl:523-536): unchanged when it has at most prod(shape)
    points or a single moment, or when the fit fails."""
    points = np.asarray(points, dtype=np.float64)
    weights = np.asarray(weights, dtype=np.float64)
    min_points = int(np.prod(shape))
    if len(points) <= min_points:
        return points, weights
    if int(np.prod([q + 1 for q in shape])) <= 1:
        return points, weights
    A = moment_matrix(points, shape)
    moments = A @ weights
    red = point_elimination(A, moments, min_points, tolerance())
    if red is None:
        return points, weights
    idx, w = red
    return points[idx], np.asarray(w)
