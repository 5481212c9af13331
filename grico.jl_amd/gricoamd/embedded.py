"""Finite-cell geometry on the host: implicit regions, cut-cell quadratures and
embedded-surface quadratures for a sphere.

Reference: src/regions.jl. Leaves and subcells are classified by the sign of
the classifier at their corners and midpoint (_classify_subcell 736-757:
negative = inside); cut cells get a candidate rule by recursive bisection to
`subdivision_depth` (_collect_finite_cell_candidates_recursive! 477-517):
inside subcells take the cell's tensor rule mapped to the subcell, terminal
cut subcells keep only their inside points, with the corner/midpoint fallback
for slivers (759-817). The reference then compresses the candidates by
non-negative moment fitting (519-708, NNLS with SVD); that step is not
restated here: the candidate rule is used as is, so cut cells carry more
points than in the reference, and the points and weights of cut cells are
parity unpinned (the reference's own tests check them only for finiteness).

The surface rule for a sphere is a product rule on the sphere (Gauss-Legendre
in cos(theta), uniform in phi) whose points are assigned to the leaves that
contain them and mapped to reference coordinates with the weight and normal
transforms of src/integration.jl:2501-2516.
"""

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


def add_cut_cell_quadratures(problem, region, cut_leaves, alpha=0.0):
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
        for leaf, rule in cut_cell_rules(space.domain, region, ls, shape).items():
            if alpha > 0.0:
                rule = PointQuadrature(np.concatenate([rule.points, fp], axis=0),
                                       np.concatenate([(1.0 - alpha) * rule.weights, alpha * fw]))
            problem.add_cell_quadrature(leaf, rule)
            total += rule.point_count()
    return total


def sphere_surface_quadratures(space, center, radius, n_theta=64, n_phi=128):
    """SurfaceQuadrature pieces of the sphere |x - center| = radius on the
    leaves of a space whose leaves all have the same size (uniform meshes);
    outward normals."""
    if space.dim != 3:
        raise ValueError("the sphere rule is three-dimensional")
    dom = space.domain
    mu, wmu = gauss_legendre_rule(n_theta)
    phi = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
    M, PH = np.meshgrid(mu, phi, indexing="ij")
    W = np.outer(wmu, np.full(n_phi, 2.0 * np.pi / n_phi)) * radius ** 2
    st = np.sqrt(1.0 - M ** 2)
    nrm = np.stack([st * np.cos(PH), st * np.sin(PH), M], axis=-1).reshape(-1, 3)
    x = np.asarray(center, dtype=np.float64)[None, :] + radius * nrm
    w = W.reshape(-1)
    sizes = dom.cell_sizes(space.leaves)
    if not np.allclose(sizes, sizes[0]):
        raise ValueError("sphere_surface_quadratures needs leaves of one size")
    h = sizes[0]
    lowers = dom.cell_lowers(space.leaves)
    origin = np.asarray(dom.origin, dtype=np.float64)
    key = {tuple(np.round((lo - origin) / h).astype(np.int64)): int(l) for lo, l in zip(lowers, space.leaves)}
    ijk = np.floor((x - origin) / h).astype(np.int64)
    groups = {}
    for p in range(x.shape[0]):
        leaf = key.get(tuple(ijk[p]))
        if leaf is not None:
            groups.setdefault(leaf, []).append(p)
    out = []
    half = 0.5 * h
    for leaf, ps in groups.items():
        ps = np.asarray(ps)
        lo = dom.cell_lowers([leaf])[0]
        xi = np.clip((x[ps] - lo) / half - 1.0, -1.0, 1.0)
        n_phys = nrm[ps]
        # reference normal: J n normalized; reference weight: physical / (det J / |J n_ref|)
        nr = n_phys * half
        nr = nr / np.linalg.norm(nr, axis=1)[:, None]
        scale = np.prod(half) / np.linalg.norm(half * nr, axis=1)
        out.append(SurfaceQuadrature(leaf, PointQuadrature(xi, w[ps] / scale), nr))
    return out
