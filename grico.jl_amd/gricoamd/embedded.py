"""Finite-cell geometry on the host: implicit regions, cut-cell quadratures and
embedded-surface quadratures for a sphere.

Reference: src/regions.jl. Leaves and subcells are classified by the sign of
the classifier at their corners and midpoint (_classify_subcell 736-757:
negative = inside); cut cells get a candidate rule by recursive bisection to
`subdivision_depth` (_collect_finite_cell_candidates_recursive! 477-517):
inside subcells take the cell's tensor rule mapped to the subcell, terminal
cut subcells keep only their inside points, with the corner/midpoint fallback
for slivers (759-817). The candidates are then compressed by the reference's
non-negative moment fitting (519-708): the Legendre moments of the full
tensor box 0..q_a, the NNLS active-set loop with SVD minimum-norm solves, and
point elimination down to prod(q) points while the moments stay fitted to
1000 eps. Where the fit does not reach that tolerance the candidate rule is
kept, as in the reference (on the 48^3 unit-ball mesh at depth 3 about half
of the cut cells keep their candidates and the fitted ones stop at
(q + 1)^D points, where no single removal keeps the fit).

The surface rule for a sphere is a product rule on the sphere (Gauss-Legendre
in cos(theta), uniform in phi) whose points are assigned to the leaves that
contain them and mapped to reference coordinates with the weight and normal
transforms of src/integration.jl:2501-2516.
"""

import os

import numpy as np

from .problem import PointQuadrature, SurfaceQuadrature
from .quadrature import gauss_legendre_rule
from .space import HpSpace

INSIDE, OUTSIDE, CUT = 0, 1, 2


class ImplicitRegion:
    """Region {x : classifier(x) < 0} (src/regions.jl:177-196). `classifier`
    takes a tuple of coordinate arrays and returns an array (or booleans,
    True = inside)."""

    def __init__(self, classifier, subdivision_depth=2):
        d = int(subdivision_depth)
        if d < 0:
            raise ValueError("subdivision_depth must be non-negative")
        self.classifier = classifier
        self.subdivision_depth = d

    def value(self, x):
        v = self.classifier(tuple(x))
        v = np.asarray(v)
        if v.dtype == bool:
            return np.where(v, -1.0, 1.0)
        v = np.asarray(v, dtype=np.float64)
        if not np.all(np.isfinite(v)):
            raise ValueError("finite-cell classifier values must be finite")
        return v


def _corner_offsets(D):
    return np.array([[(m >> a) & 1 for a in range(D)] for m in range(1 << D)], dtype=np.float64)


def _to_physical(lowers, halves, xi):
    """Physical coordinates of reference points xi (..., D) in cells with
    lower corners / half sizes (cells, D)."""
    return lowers[..., :] + halves * (xi + 1.0)


def classify(domain, region, cells, lower, upper):
    """INSIDE / OUTSIDE / CUT of subcells [lower, upper] (biunit coordinates,
    arrays (n, D)) of `cells` (n) by corner and midpoint samples."""
    D = domain.dim
    lo_c = domain.cell_lowers(cells)
    half = 0.5 * domain.cell_sizes(cells)
    corners = _corner_offsets(D)
    samples = [lower + corners[m] * (upper - lower) for m in range(len(corners))]
    samples.append(0.5 * (lower + upper))
    neg = np.zeros(len(cells), dtype=bool)
    pos = np.zeros(len(cells), dtype=bool)
    for xi in samples:
        x = _to_physical(lo_c, half, xi)
        v = region.value([x[:, a] for a in range(D)])
        neg |= v < 0
        pos |= v > 0
    out = np.full(len(cells), CUT, dtype=np.int8)
    out[~neg] = OUTSIDE
    out[neg & ~pos] = INSIDE
    return out


def classify_leaves(domain, region, leaves):
    leaves = np.asarray(leaves, dtype=np.int64)
    D = domain.dim
    lo = -np.ones((len(leaves), D))
    return classify(domain, region, leaves, lo, -lo)


def cut_cell_rules(domain, region, leaves, quadrature_shape):
    """Candidate quadratures (reference coordinates) of cut `leaves`: dict leaf
    -> PointQuadrature (src/regions.jl:304-329 without the moment fit)."""
    leaves = np.asarray(leaves, dtype=np.int64)
    D = domain.dim
    rules = [gauss_legendre_rule(n) for n in quadrature_shape]
    grids = np.meshgrid(*[r[0] for r in rules][::-1], indexing="ij")
    eta = np.stack([gg.reshape(-1) for gg in grids[::-1]], axis=1)
    wg = np.meshgrid(*[r[1] for r in rules][::-1], indexing="ij")
    w_eta = np.ones_like(wg[0])
    for gg in wg:
        w_eta = w_eta * gg
    w_eta = w_eta.reshape(-1)
    corners = _corner_offsets(D)
    pts = {int(l): [] for l in leaves}
    wts = {int(l): [] for l in leaves}
    cell = leaves.copy()
    lower = -np.ones((len(leaves), D))
    upper = np.ones((len(leaves), D))
    lo_all = {int(l): v for l, v in zip(leaves, domain.cell_lowers(leaves))}
    half_all = {int(l): v for l, v in zip(leaves, 0.5 * domain.cell_sizes(leaves))}
    for depth in range(region.subdivision_depth + 1):
        if len(cell) == 0:
            break
        cls = classify(domain, region, cell, lower, upper)
        terminal = depth >= region.subdivision_depth
        emit = (cls == INSIDE) | ((cls == CUT) & terminal)
        for k in np.nonzero(emit)[0]:
            c = int(cell[k])
            center = 0.5 * (lower[k] + upper[k])
            hs = 0.5 * (upper[k] - lower[k])
            xi = center + eta * hs
            w = w_eta * np.prod(hs)
            if cls[k] == CUT:
                x = _to_physical(lo_all[c], half_all[c], xi)
                keep = region.value([x[:, a] for a in range(D)]) < 0
                if not np.any(keep):
                    # sliver fallback: the negative corner / midpoint samples
                    sw = np.prod(upper[k] - lower[k]) / (len(corners) + 1)
                    samp = np.vstack([center[None, :], lower[k] + corners * (upper[k] - lower[k])])
                    xs = _to_physical(lo_all[c], half_all[c], samp)
                    neg = region.value([xs[:, a] for a in range(D)]) < 0
                    xi, w = samp[neg], np.full(int(neg.sum()), sw)
                else:
                    xi, w = xi[keep], w[keep]
            pts[c].append(xi)
            wts[c].append(w)
        split = np.nonzero((cls == CUT) & (not terminal))[0]
        if len(split) == 0:
            break
        mid = 0.5 * (lower[split] + upper[split])
        new_cell, new_lo, new_up = [], [], []
        for m in range(1 << D):
            bits = corners[m].astype(bool)
            new_cell.append(cell[split])
            new_lo.append(np.where(bits, mid, lower[split]))
            new_up.append(np.where(bits, upper[split], mid))
        cell = np.concatenate(new_cell)
        lower = np.concatenate(new_lo)
        upper = np.concatenate(new_up)
    out = {}
    for c in pts:
        if pts[c]:
            out[c] = PointQuadrature(np.concatenate(pts[c], axis=0), np.concatenate(wts[c]))
    return out


# ---- non-negative moment fitting (src/regions.jl:519-708) ----------------------

_MOMENT_TOLERANCE = 1.0e3 * np.finfo(np.float64).eps  # 519, 523
_WEIGHT_PRUNE_RATIO = 1.0e-8                          # 520
_MAX_ELIMINATION_ITERATIONS = 1024                    # 521


def _legendre_table(x, degree):
    """P_0..P_degree at the points x (n) by the three-term recurrence
    (src/polynomials.jl:294-307): (n, degree + 1)."""
    t = np.empty((len(x), degree + 1))
    t[:, 0] = 1.0
    if degree >= 1:
        t[:, 1] = x
    for n in range(2, degree + 1):
        inv_n = 1.0 / n
        t[:, n] = (2 * n - 1) * inv_n * x * t[:, n - 1] - (n - 1) * inv_n * t[:, n - 2]
    return t


def moment_matrix(points, shape):
    """Legendre moments of the full tensor box 0..q_a at the points: modes
    (first axis fastest) x points (src/regions.jl:538-561), each product
    formed axis by axis in the reference's order (rounding-identical)."""
    n, D = points.shape
    grids = np.meshgrid(*[np.arange(q + 1) for q in shape][::-1], indexing="ij")
    modes = [g.reshape(-1) for g in grids[::-1]]  # per axis, first axis fastest
    M = np.ones((len(modes[0]), n))
    for a in range(D):
        M = M * _legendre_table(points[:, a], shape[a]).T[modes[a]]
    return M


def _least_squares_min_norm(A, b, tol):
    """Minimum-norm least squares, singular values at most tol * s_max
    dropped (src/regions.jl:709-725)."""
    if A.shape[1] == 0:
        return np.zeros(0)
    U, S, Vt = np.linalg.svd(A, full_matrices=False)
    keep = S > tol * S[0]
    scaled = np.where(keep, (U.T @ b) / np.where(keep, S, 1.0), 0.0)
    return Vt.T @ scaled


def _nnls(A, b, tol):
    """Non-negative least squares by the reference's active-set loop
    (src/regions.jl:641-707): the entering column is the first one of largest
    positive gradient, passive solves are minimum-norm, infeasible steps are
    cut back to the first passive weight that reaches tol. Returns (weights,
    |b - A w| / max(|b|, 1))."""
    n = A.shape[1]
    w = np.zeros(n)
    passive = np.zeros(n, dtype=bool)
    gradient = A.T @ b
    for _ in range(max(8 * n, 32)):
        g = np.where(passive, 0.0, gradient)
        cand = int(np.argmax(g))
        if not g[cand] > 0.0 or not gradient[cand] > tol:
            break
        passive[cand] = True
        while True:
            idx = np.nonzero(passive)[0]
            sol = _least_squares_min_norm(A[:, idx], b, tol)
            if np.all(sol > tol):
                w = np.zeros(n)
                w[idx] = sol
                break
            cur = w[idx]
            den = cur - sol
            step = (sol <= tol) & (np.abs(den) > tol)
            alpha = min(1.0, float(np.min(cur[step] / den[step]))) if np.any(step) else 1.0
            w[idx] = cur + alpha * (sol - cur)
            drop = idx[w[idx] <= tol]
            w[drop] = 0.0
            passive[drop] = False
            if not np.any(passive):
                break
        gradient = A.T @ (b - A @ w)
    return w, float(np.linalg.norm(b - A @ w) / max(np.linalg.norm(b), 1.0))


def _removable(weights, min_points, tol):
    """src/regions.jl:621-630: weights at most max(tol, 1e-8 max w) in
    increasing order, else the smallest one."""
    if len(weights) <= min_points:
        return []
    threshold = max(tol, _WEIGHT_PRUNE_RATIO * float(np.max(weights)))
    small = np.nonzero(weights <= threshold)[0]
    if len(small) == 0:
        return [int(np.argmin(weights))]
    return [int(i) for i in small[np.argsort(weights[small], kind="stable")]]


def moment_fit(points, weights, shape):
    """Fitted rule (points, weights) of a candidate rule on the biunit cell
    (src/regions.jl:523-536, _point_elimination 563-619); the candidate itself
    when it has at most prod(shape) points or the fit misses the tolerance."""
    points = np.asarray(points, dtype=np.float64)
    weights = np.asarray(weights, dtype=np.float64)
    min_points = int(np.prod(shape))
    if len(points) <= min_points or int(np.prod([q + 1 for q in shape])) <= 1:
        return points, weights
    tol = _MOMENT_TOLERANCE
    A = moment_matrix(points, shape)
    moments = A @ weights
    w, res = _nnls(A, moments, tol)
    if not res <= tol:
        return points, weights
    current = np.nonzero(w > tol)[0]
    if len(current) == 0:
        return points, weights
    cw = w[current]
    if len(current) > A.shape[0]:
        current = current[np.argsort(-cw, kind="stable")[:A.shape[0]]]
        cw, res = _nnls(A[:, current], moments, tol)
        if not res <= tol:
            return points, weights
    for _ in range(_MAX_ELIMINATION_ITERATIONS):
        if len(current) <= min_points:
            break
        accepted = False
        for li in _removable(cw, min_points, tol):
            trial = np.delete(current, li)
            if len(trial) < min_points:
                continue
            tw, res = _nnls(A[:, trial], moments, tol)
            if not res <= tol:
                continue
            current, cw, accepted = trial, tw, True
            break
        if not accepted:
            break
    return points[current], cw


def _fit_rules(items):
    return [moment_fit(p, w, shape) for p, w, shape in items]


def fitted_cut_cell_rules(domain, region, leaves, quadrature_shape, workers=None, progress=None):
    """cut_cell_rules compressed by moment fitting (src/regions.jl:304-329):
    dict leaf -> PointQuadrature. Cells are fitted in `workers` spawned
    processes (default: the host's CPUs, at most 16) when there are many;
    `progress(done, total)` is called as batches of cells finish."""
    cands = cut_cell_rules(domain, region, leaves, quadrature_shape)
    keys = list(cands)
    items = [(cands[k].points, cands[k].weights, tuple(quadrature_shape)) for k in keys]
    if workers is None:
        workers = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        workers = max(1, min(int(workers), 16))
    if workers > 1 and len(items) >= 64:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        chunk = max(1, len(items) // (8 * workers))
        parts = [items[i:i + chunk] for i in range(0, len(items), chunk)]
        # one BLAS thread per worker: the solves are small and many, and
        # threaded BLAS in every worker oversubscribes the cores badly
        names = ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS")
        saved = {k: os.environ.get(k) for k in names}
        os.environ.update({k: "1" for k in names})
        try:
            with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
                fitted = []
                for part in ex.map(_fit_rules, parts):
                    fitted.extend(part)
                    if progress is not None:
                        progress(len(fitted), len(items))
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
    else:
        fitted = _fit_rules(items)
    return {k: PointQuadrature(p, w) for k, (p, w) in zip(keys, fitted)}


# ---- physical domains (src/regions.jl:40-184, 255-395) ---------------------------

class PhysicalMeasure:
    """Cut-cell integrals over the physical part of the cell only
    (src/regions.jl:45-53)."""


class FiniteCellExtension:
    """Cut-cell measure (1 - alpha) Omega_phys + alpha Omega_full
    (src/regions.jl:55-80): the fictitious part of a cut cell keeps weight
    alpha, which bounds the conditioning of small cuts."""

    def __init__(self, alpha):
        alpha = float(alpha)
        if not np.isfinite(alpha):
            raise ValueError("finite-cell extension alpha must be finite")
        if not 0.0 <= alpha <= 1.0:
            raise ValueError("finite-cell extension alpha must lie in [0, 1]")
        self.alpha = alpha


class PhysicalDomain:
    """A background Domain together with an implicit region and a cut-cell
    measure (src/regions.jl:125-184). Spaces built on it drop the leaves
    classified outside (255-300); compile() gives every cut leaf of the space
    its moment-fitted rule for the leaf's quadrature shape (304-329, cached
    per leaf and shape like the reference's signature cache) with the cell
    measure applied (338-380). The background's geometry and refinement tree
    are used as they are."""

    def __init__(self, background, region, cell_measure=None):
        if isinstance(background, PhysicalDomain):
            raise TypeError("PhysicalDomain wraps a background Domain")
        if not isinstance(region, ImplicitRegion):
            raise TypeError("region must be an ImplicitRegion")
        measure = PhysicalMeasure() if cell_measure is None else cell_measure
        if not isinstance(measure, (PhysicalMeasure, FiniteCellExtension)):
            raise TypeError("cell_measure must be PhysicalMeasure() or FiniteCellExtension(alpha)")
        self.background = background
        self.region = region
        self.cell_measure = measure
        self._fitted = {}  # (leaf, shape) -> fitted physical rule, or None
        self.fit_progress = None  # optional progress(done, total) of the moment fits

    def __getattr__(self, name):
        if name.startswith("__") or name == "background":
            raise AttributeError(name)
        return getattr(self.background, name)

    def classify(self, leaves):
        return classify_leaves(self.background, self.region, leaves)

    def active_leaves(self):
        leaves = np.asarray(self.background.active_leaves(), dtype=np.int64)
        keep = leaves[self.classify(leaves) != OUTSIDE]
        if len(keep) == 0:
            raise ValueError("physical domains must contain at least one active physical leaf")
        return keep

    def physical_rules(self, leaves, shape, progress=None):
        """Moment-fitted physical rules of cut `leaves` for one quadrature
        shape (dict leaf -> PointQuadrature; leaves whose candidate rule is
        empty are absent)."""
        shape = tuple(int(q) for q in shape)
        todo = [int(l) for l in leaves if (int(l), shape) not in self._fitted]
        if todo:
            got = fitted_cut_cell_rules(self.background, self.region, todo, shape,
                                        progress=progress if progress is not None else self.fit_progress)
            for l in todo:
                self._fitted[(l, shape)] = got.get(l)
        return {int(l): self._fitted[(int(l), shape)] for l in leaves if self._fitted[(int(l), shape)] is not None}

    def assembly_cell_quadratures(self, space, progress=None):
        """leaf -> PointQuadrature of every cut leaf of `space` under the cell
        measure (src/integration.jl:2406-2418, src/regions.jl:338-380)."""
        measure = self.cell_measure
        if isinstance(measure, FiniteCellExtension) and measure.alpha == 1.0:
            return {}
        cls = self.classify(space.leaves)
        shapes = {}
        for li in np.nonzero(cls == CUT)[0]:
            shapes.setdefault(tuple(int(q) for q in space.quadrature_shapes[li]), []).append(int(space.leaves[li]))
        out = {}
        for shape, ls in shapes.items():
            rules = self.physical_rules(ls, shape, progress=progress)
            if isinstance(measure, FiniteCellExtension) and measure.alpha > 0.0:
                fp, fw = _tensor_rule(shape)
                a = measure.alpha
                rules = {l: PointQuadrature(np.concatenate([r.points, fp], axis=0),
                                            np.concatenate([(1.0 - a) * r.weights, a * fw]))
                         for l, r in rules.items()}
            out.update(rules)
        return out


class SampledSurface:
    """An embedded surface given by physical quadrature points, weights and
    outward normals, converted at compile time into SurfaceQuadrature pieces
    on whichever leaves of the space contain the points (the reference's
    EmbeddedSurface role, src/embedded.jl:140-187, for 3D surfaces the
    reference only takes as explicit per-leaf pieces: the pieces then follow
    the leaves of every multigrid level). Reference weights and normals use
    the transforms of src/integration.jl:2501-2516."""

    def __init__(self, points, weights, normals):
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        self.normals = np.ascontiguousarray(normals, dtype=np.float64)
        n = len(self.weights)
        if self.points.ndim != 2 or self.points.shape[0] != n or self.normals.shape != self.points.shape:
            raise ValueError("one point, weight and normal per surface sample is required")

    def surface_quadratures(self, space):
        dom = space.domain
        D = space.dim
        if self.points.shape[1] != D:
            raise ValueError("surface dimension must match the space")
        owner = locate_points(dom, space.leaves, self.points)
        found = np.nonzero(owner >= 0)[0]
        # pieces in the order of their first sample
        first = {}
        for p_ in found:
            first.setdefault(int(owner[p_]), []).append(p_)
        out = []
        lowers = dom.cell_lowers(space.leaves)
        halves = 0.5 * dom.cell_sizes(space.leaves)
        for li, ps in first.items():
            ps = np.asarray(ps)
            lo, half = lowers[li], halves[li]
            xi = np.clip((self.points[ps] - lo) / half - 1.0, -1.0, 1.0)
            nr = self.normals[ps] * half
            nr = nr / np.linalg.norm(nr, axis=1)[:, None]
            scale = np.prod(half) / np.linalg.norm(half * nr, axis=1)
            out.append(SurfaceQuadrature(int(space.leaves[li]), PointQuadrature(xi, self.weights[ps] / scale), nr))
        return out


def locate_points(domain, leaves, x):
    """Index into `leaves` of the leaf containing each point of x (n, D), or -1:
    leaves are grouped by size, and each group is looked up by the integer
    coordinates of the points on that group's lattice."""
    leaves = np.asarray(leaves, dtype=np.int64)
    x = np.asarray(x, dtype=np.float64)
    out = np.full(len(x), -1, dtype=np.int64)
    if len(leaves) == 0 or len(x) == 0:
        return out
    origin = np.asarray(domain.origin, dtype=np.float64)
    extent = np.asarray(domain.extent, dtype=np.float64)
    sizes = domain.cell_sizes(leaves)
    lowers = domain.cell_lowers(leaves)
    groups = {}
    for i, h in enumerate(map(tuple, sizes)):
        groups.setdefault(h, []).append(i)
    for h, idx in groups.items():
        idx = np.asarray(idx)
        h = np.asarray(h)
        k = np.rint((lowers[idx] - origin) / h).astype(np.int64)
        ext = k.max(axis=0) + 2
        code = np.zeros(len(idx), dtype=np.int64)
        for a in range(len(h)):
            code = code * ext[a] + k[:, a]
        order = np.argsort(code)
        codes = code[order]
        kp = np.floor((x - origin) / h).astype(np.int64)
        # a point on the domain's upper face belongs to the last cell of the
        # lattice at this size (interior faces belong to the upper cell)
        kp = np.minimum(kp, np.rint(extent / h).astype(np.int64) - 1)
        ok = np.all(kp >= 0, axis=1) & (out < 0)
        pc = np.zeros(len(x), dtype=np.int64)
        for a in range(len(h)):
            pc = pc * ext[a] + np.clip(kp[:, a], 0, ext[a] - 1)
        pos = np.searchsorted(codes, pc)
        pos = np.minimum(pos, len(codes) - 1)
        hit = ok & (codes[pos] == pc)
        out[hit] = idx[order[pos[hit]]]
    return out


def sphere_surface(center, radius, n_theta=64, n_phi=128):
    """Samples of the sphere |x - center| = radius: Gauss-Legendre in
    cos(theta) times uniform in phi, outward normals."""
    mu, wmu = gauss_legendre_rule(n_theta)
    phi = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
    M, PH = np.meshgrid(mu, phi, indexing="ij")
    W = np.outer(wmu, np.full(n_phi, 2.0 * np.pi / n_phi)) * radius ** 2
    st = np.sqrt(1.0 - M ** 2)
    nrm = np.stack([st * np.cos(PH), st * np.sin(PH), M], axis=-1).reshape(-1, 3)
    x = np.asarray(center, dtype=np.float64)[None, :] + radius * nrm
    return SampledSurface(x, W.reshape(-1), nrm)


def physical_space(domain, region, options):
    """HpSpace on the leaves that are not outside the region
    (src/regions.jl:267-288), plus their classification."""
    leaves = domain.active_leaves()
    cls = classify_leaves(domain, region, leaves)
    active = leaves[cls != OUTSIDE]
    if len(active) == 0:
        raise ValueError("physical domains must contain at least one active physical leaf")
    space = HpSpace(domain, options, leaves=active)
    return space, active[cls[cls != OUTSIDE] == CUT]


def _tensor_rule(shape):
    rules = [gauss_legendre_rule(n) for n in shape]
    grids = np.meshgrid(*[r[0] for r in rules][::-1], indexing="ij")
    pts = np.stack([gg.reshape(-1) for gg in grids[::-1]], axis=1)
    wg = np.meshgrid(*[r[1] for r in rules][::-1], indexing="ij")
    w = np.ones_like(wg[0])
    for gg in wg:
        w = w * gg
    return pts, w.reshape(-1)


def add_cut_cell_quadratures(problem, region, cut_leaves, alpha=0.0, progress=None):
    """Attach the cut-cell rules of `cut_leaves` to the problem; returns the
    total number of cut-cell points. alpha > 0 is the finite-cell extension
    (src/regions.jl:56-87, 352-380): physical weights scaled by 1 - alpha plus
    alpha times the full tensor rule of the cell."""
    alpha = float(alpha)
    if not (0.0 <= alpha <= 1.0):
        raise ValueError("alpha must lie in [0, 1]")
    space = problem.field.space
    shapes = {}
    for leaf in cut_leaves:
        shapes.setdefault(space.cell_quadrature_shape(int(leaf)), []).append(int(leaf))
    total = 0
    for shape, ls in shapes.items():
        fp, fw = _tensor_rule(shape)
        for leaf, rule in fitted_cut_cell_rules(space.domain, region, ls, shape, progress=progress).items():
            if alpha > 0.0:
                rule = PointQuadrature(np.concatenate([rule.points, fp], axis=0),
                                       np.concatenate([(1.0 - alpha) * rule.weights, alpha * fw]))
            problem.add_cell_quadrature(leaf, rule)
            total += rule.point_count()
    return total


def sphere_surface_quadratures(space, center, radius, n_theta=64, n_phi=128):
    """SurfaceQuadrature pieces of the sphere |x - center| = radius on the
    leaves of `space` (outward normals)."""
    if space.dim != 3:
        raise ValueError("the sphere rule is three-dimensional")
    return sphere_surface(center, radius, n_theta, n_phi).surface_quadratures(space)
