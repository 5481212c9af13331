"""Problem builders shared by the tests (mirroring the reference's test problems
and the five benchmark configurations at reduced sizes)."""

import numpy as np

import gricoamd as g


class Diffusion:
    """TestChannels(0, k grad u) (examples/poisson_matrix_free_gmg/driver.jl:34-36)."""

    def __init__(self, kappa=None):
        self.kappa = kappa

    def cell_accumulate(self, q, trial, test_component):
        k = 1.0 if self.kappa is None else self.kappa(q.point)
        return g.TestChannels(0.0 * trial.value, tuple(k * c for c in trial.gradient))


class Mass:
    def __init__(self, scale=1.0):
        self.scale = scale

    def cell_accumulate(self, q, trial, test_component):
        return self.scale * trial.value


class MassDiffusion:
    """Nonsymmetric-capable general form: a0 = m u + b.grad u, a_grad = k grad u."""

    def __init__(self, m=1.0, k=1.0, b=None):
        self.m, self.k, self.b = m, k, b

    def cell_accumulate(self, q, trial, test_component):
        v = self.m * trial.value
        if self.b is not None:
            for a, ba in enumerate(self.b):
                v = v + ba * trial.gradient[a]
        return g.TestChannels(v, tuple(self.k * c for c in trial.gradient))


class SineLoad:
    def __init__(self, dim):
        self.dim = dim

    def cell_rhs_accumulate(self, q, test_component):
        out = self.dim * np.pi ** 2
        for x in q.point:
            out = out * np.sin(np.pi * x)
        return out


class CosLoad:
    """f = D pi^2 prod cos(pi x_a): with zero normal flux and zero mean the
    solution is prod cos(pi x_a)."""

    def __init__(self, dim):
        self.dim = dim

    def cell_rhs_accumulate(self, q, test_component):
        out = self.dim * np.pi ** 2
        for x in q.point:
            out = out * np.cos(np.pi * x)
        return out


def cos_exact(p):
    out = 1.0
    for x in p:
        out = out * np.cos(np.pi * x)
    return out


def mean_golden_problem():
    """One DG p = 0 cell on [0, 1] with an integral-mean constraint 2
    (test/matrix_free.jl:1047-1053)."""
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(0), continuity="dg"))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Mass())
    problem.add_constraint(g.MeanValue(u, 2.0))
    return problem


def sine_exact(p):
    out = 1.0
    for x in p:
        out = out * np.sin(np.pi * x)
    return out


def kappa_c2(p):
    """kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z)."""
    return 1.0 + 0.5 * np.sin(2 * np.pi * p[0]) * np.sin(2 * np.pi * p[1]) * np.sin(2 * np.pi * p[2])


def poisson(dim, cells, degree, kappa=None, dirichlet=True, basis=None, rhs=True, refine_depth=0,
            operator=None, continuity="cg"):
    if refine_depth:
        domain = g.Domain((0.0,) * dim, (1.0,) * dim, (1,) * dim)
        g.uniform_refine(domain, refine_depth)
    else:
        domain = g.Domain((0.0,) * dim, (1.0,) * dim, (cells,) * dim if np.ndim(cells) == 0 else cells)
    space = g.HpSpace(domain, g.SpaceOptions(basis=basis or g.FullTensorBasis(), degree=g.UniformDegree(degree),
                                             continuity=continuity))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, operator or Diffusion(kappa))
    if rhs:
        problem.add_cell_accumulator(u, SineLoad(dim))
    if dirichlet:
        for axis in range(dim):
            for side in (g.LOWER, g.UPPER):
                problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return problem


def hanging_problem(degree=2, operator=None, dim=2):
    """Two root cells, the second bisected: one hanging face (test/space.jl:304-316)."""
    if dim == 2:
        domain = g.Domain((0.0, 0.0), (2.0, 1.0), (2, 1))
        domain.refine(1, 1)
    else:
        domain = g.Domain((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (2, 1, 1))
        first = domain.refine(1, 1)
        domain.refine(first, 2)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, operator or MassDiffusion(1.0, 1.0))
    return problem


def corner_refined(dim=2, base=4, levels=2, degree=2, operator=None, dirichlet=True):
    """Uniform base grid with repeated all-axis refinement of the cell at the origin."""
    domain = g.Domain((0.0,) * dim, (1.0,) * dim, (base,) * dim)
    target = 0
    for _ in range(levels):
        leaves = [target]
        for axis in range(dim):
            nxt = []
            for c in leaves:
                f = domain.refine(c, axis)
                nxt += [f, f + 1]
            leaves = nxt
        target = leaves[0]
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, operator or Diffusion())
    problem.add_cell_accumulator(u, SineLoad(dim))
    if dirichlet:
        for axis in range(dim):
            for side in (g.LOWER, g.UPPER):
                problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return problem


def rng_vector(n, seed=0):
    return np.random.default_rng(seed).standard_normal(n)


class LogNormalKappa:
    """kappa = exp(g), g a Gaussian random field with squared-exponential
    covariance sigma^2 exp(-|x - y|^2 / (2 l^2)), sampled by random Fourier
    features: g(x) = sigma sqrt(2 / M) sum_m cos(w_m . x + phi_m), w_m ~ N(0, I / l^2),
    phi_m ~ U(0, 2 pi) (seeded)."""

    def __init__(self, seed=0, corr=0.1, sigma=1.0, features=256, dim=3):
        rng = np.random.default_rng(seed)
        self.w = rng.standard_normal((features, dim)) / corr
        self.phi = rng.uniform(0.0, 2.0 * np.pi, size=features)
        self.scale = sigma * np.sqrt(2.0 / features)

    def __call__(self, p):
        p = [np.asarray(c, dtype=np.float64) for c in p]
        shape = p[0].shape
        x = np.stack([c.reshape(-1) for c in p], axis=1)
        g = self.scale * np.cos(x @ self.w.T + self.phi).sum(axis=1)
        return np.exp(g).reshape(shape)


def lognormal_poisson(depth=2, degree=3, seed=0, dim=3):
    """The random-coefficient Poisson problem (p = 3: 64 local modes per cell)."""
    return poisson(dim, None, degree, refine_depth=depth, kappa=LogNormalKappa(seed, dim=dim))


class BallLoad:
    """-lap(1 - r^2) = 2 D."""

    def __init__(self, dim=3):
        self.dim = dim

    def cell_rhs_accumulate(self, q, test_component):
        return 2.0 * self.dim + 0.0 * q.point[0]


class NitscheData:
    """Data terms of symmetric Nitsche for u = g on the embedded surface:
    g (gamma / h v - dv/dn), with g a smooth function of the point (value and
    gradient test channels both nonzero)."""

    def __init__(self, gamma):
        self.gamma = gamma

    def surface_rhs_accumulate(self, q, test_component):
        x = q.point
        gval = 1.0 + 0.5 * x[0] - 0.25 * x[1] * x[2]
        return g.TestChannels(self.gamma / q.cell_size[0] * gval, tuple(-gval * nc for nc in q.normal))


class NitscheDirichlet:
    """Symmetric Nitsche terms for u = 0 on the embedded surface:
    -(du/dn) v - u (dv/dn) + gamma / h u v."""

    def __init__(self, gamma):
        self.gamma = gamma

    def surface_accumulate(self, q, trial, test_component):
        n = q.normal
        dn = sum(nc * gc for nc, gc in zip(n, trial.gradient))
        return g.TestChannels(-dn + self.gamma / q.cell_size[0] * trial.value,
                              tuple(-trial.value * nc for nc in n))


def ball_exact(p):
    return 1.0 - (p[0] ** 2 + p[1] ** 2 + p[2] ** 2)


def fcm_ball(n=8, degree=2, depth=2, gamma=None, n_theta=48, n_phi=96, alpha=1e-5):
    """Finite-cell Poisson on the unit ball in [-1.2, 1.2]^3: n^3 background
    cells, cut cells with subdivision rules, Nitsche on the sphere, exact
    solution 1 - r^2."""
    from gricoamd.embedded import ImplicitRegion, add_cut_cell_quadratures, physical_space, \
        sphere_surface_quadratures
    region = ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=depth)
    dom = g.Domain((-1.2, -1.2, -1.2), (2.4, 2.4, 2.4), (n, n, n))
    space, cut = physical_space(dom, region, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, Diffusion())
    pr.add_cell_accumulator(u, BallLoad(3))
    add_cut_cell_quadratures(pr, region, cut, alpha=alpha)
    for sq in sphere_surface_quadratures(space, (0.0, 0.0, 0.0), 1.0, n_theta=n_theta, n_phi=n_phi):
        pr.add_surface_quadrature(sq)
    pr.add_surface_accumulator(u, u, NitscheDirichlet(gamma if gamma is not None else 10.0 * (degree + 1) ** 2))
    return pr


def _ball_background(n):
    """[-1.2, 1.2]^3 with n^3 leaves built as roots^3 root cells refined k
    times (n = roots 2^k, roots odd), so that multigrid finds h-levels."""
    roots, k = n, 0
    while roots % 2 == 0:
        roots, k = roots // 2, k + 1
    dom = g.Domain((-1.2, -1.2, -1.2), (2.4, 2.4, 2.4), (roots,) * 3)
    g.uniform_refine(dom, k)
    return dom


def fcm_ball_physical(n=4, degree=2, depth=2, gamma=None, n_theta=48, n_phi=96, alpha=1e-3, explicit=False):
    """fcm_ball on a PhysicalDomain: the cut-cell rules are generated by
    compile() for whatever leaves the space has (moment fitted, finite-cell
    extension alpha) and the sphere is a SampledSurface; explicit=True attaches
    the same rules and pieces per leaf instead (add_cell_quadrature /
    add_surface_quadrature)."""
    from gricoamd.embedded import (FiniteCellExtension, ImplicitRegion, PhysicalDomain, add_cut_cell_quadratures,
                                   physical_space, sphere_surface)
    region = ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=depth)
    opts = g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree))
    surface = sphere_surface((0.0, 0.0, 0.0), 1.0, n_theta=n_theta, n_phi=n_phi)
    if explicit:
        space, cut = physical_space(_ball_background(n), region, opts)
    else:
        space = g.HpSpace(PhysicalDomain(_ball_background(n), region, FiniteCellExtension(alpha)), opts)
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, Diffusion())
    pr.add_cell_accumulator(u, BallLoad(3))
    if explicit:
        add_cut_cell_quadratures(pr, region, cut, alpha=alpha)
        for sq in surface.surface_quadratures(space):
            pr.add_surface_quadrature(sq)
    else:
        pr.add_embedded_surface(surface)
    pr.add_surface_accumulator(u, u, NitscheDirichlet(gamma if gamma is not None else 10.0 * (degree + 1) ** 2))
    return pr


# ---- adaptive hp corner problem ---------------------------------------------------

def corner_exact(p):
    """u = r^(2/3), r = |x| (the corner singularity at the origin)."""
    r2 = 0.0
    for x in p:
        r2 = r2 + x * x
    return np.sqrt(r2) ** (2.0 / 3.0)


class CornerLoad:
    """-div grad r^a = -a (a + 1) r^(a - 2) in 3D; a = 2/3 gives -(10/9) r^(-4/3)."""

    def cell_rhs_accumulate(self, q, test_component):
        r2 = 0.0
        for x in q.point:
            r2 = r2 + x * x
        return -(10.0 / 9.0) * r2 ** (-2.0 / 3.0)


def adaptive_hp(base=4, levels=3, pmax=8):
    """3D Poisson on [0,1]^3 shaped like the adaptive hp workload: base^3 cells
    (uniform refinements of one root cell, so that multigrid can coarsen the
    base grid too; base a power of two), `levels` rounds of dyadic refinement
    of the leaf at the corner (0,0,0)
    (all three axes on even rounds, x and y only on odd rounds, so cells are
    anisotropic and faces carry hanging nodes), per-leaf degree 1..pmax rising
    with the distance of the leaf center from the corner, u = r^(2/3)
    manufactured (load CornerLoad, Dirichlet data u on every face)."""
    depth = int(round(np.log2(base)))
    if 2 ** depth != base:
        raise ValueError("base must be a power of two")
    domain = g.Domain((0.0,) * 3, (1.0,) * 3, (1, 1, 1))
    g.uniform_refine(domain, depth)
    for lev in range(levels):
        axes = (0, 1, 2) if lev % 2 == 0 else (0, 1)
        lowers = domain.cell_lowers(domain.active_leaves())
        corner = [int(l) for l, lo in zip(domain.active_leaves(), lowers) if np.all(lo == 0.0)]
        for leaf in corner:
            cur = [leaf]
            for a in axes:
                nxt = []
                for c in cur:
                    f = domain.refine(c, a)
                    nxt += [f, f + 1]
                cur = nxt

    def degree(dom, leaf):
        lo = dom.cell_lowers([leaf])[0]
        h = dom.cell_sizes([leaf])[0]
        d = float(np.linalg.norm(lo + 0.5 * h)) / np.sqrt(3.0)
        return int(min(pmax, 1 + np.floor(pmax * d)))

    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.ByLeafDegrees(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Diffusion())
    problem.add_cell_accumulator(u, CornerLoad())
    for axis in range(3):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), corner_exact))
    return problem


# ---- symmetric interior penalty DG (benchmark/sine_poisson_sipg.jl) ----------------

def _penalty(q, degree, eta):
    """sigma = eta (p + 1)^2 / h with h the cell size across the face (the
    interface's smaller cell, src/integration.jl:2311-2313)."""
    return eta * (degree + 1) ** 2 / q.cell_size[q.face_axis]


class SIPGInterface:
    """-{d_n u}[v] - {d_n v}[u] + sigma [u][v] with [w] = w+ - w- (the reference's
    jump orientation, src/integration.jl:375); the form is invariant to the
    orientation of [.]."""

    def __init__(self, degree, eta=10.0):
        self.degree, self.eta = degree, eta

    def interface_accumulate(self, q, trial, test_component):
        n = q.normal
        jmp = g.jump(trial.value)
        avg = g.average(g.normal_component(trial.gradient, n))
        sigma = _penalty(q, self.degree, self.eta)
        shared = tuple(0.5 * jmp * c for c in n)
        return g.TraceTestChannels(g.TestChannels(-avg - sigma * jmp, shared),
                                   g.TestChannels(avg + sigma * jmp, shared))


class SIPGBoundary:
    """Weak Dirichlet (Nitsche) on a boundary face: -d_n u v - d_n v u + sigma u v."""

    def __init__(self, degree, eta=10.0):
        self.degree, self.eta = degree, eta

    def boundary_accumulate(self, q, trial, test_component):
        n = q.normal
        dn = g.normal_component(trial.gradient, n)
        sigma = _penalty(q, self.degree, self.eta)
        return g.TestChannels(-dn + sigma * trial.value, tuple(-trial.value * c for c in n))


class SIPGBoundaryLoad:
    """g (sigma v - d_n v) for Dirichlet data g."""

    def __init__(self, degree, data, eta=10.0):
        self.degree, self.data, self.eta = degree, data, eta

    def boundary_rhs_accumulate(self, q, test_component):
        gv = self.data(q.point)
        sigma = _penalty(q, self.degree, self.eta)
        return g.TestChannels(sigma * gv, tuple(-gv * c for c in q.normal))


class ConstantLoad:
    def __init__(self, value):
        self.value = value

    def cell_rhs_accumulate(self, q, test_component):
        return self.value + 0.0 * q.point[0]


def quadratic_exact(p):
    """u = x^2 + 2 y^2 - x y (+ z^2 - y z in 3D); -Laplace u = -6 (2D) / -8 (3D)."""
    out = p[0] ** 2 + 2.0 * p[1] ** 2 - p[0] * p[1]
    if len(p) > 2:
        out = out + p[2] ** 2 - p[1] * p[2]
    return out


def _refined_grid(dim, cells, refine):
    """[0,1]^dim with `cells` roots per axis and (cell, axis) bisections; cell -1
    is the first child of the previous bisection."""
    domain = g.Domain((0.0,) * dim, (1.0,) * dim, (cells,) * dim)
    last = None
    for cell, axis in refine:
        last = domain.refine(last if cell == -1 else cell, axis)
    return domain


def sipg_poisson(dim=2, cells=3, degree=2, eta=10.0, exact=None, source=None, refine=(), degrees=None,
                 rhs=True, refine_depth=0):
    """DG Poisson with SIPG interfaces and Nitsche boundaries on [0,1]^dim.
    `refine` lists (cell, axis) bisections applied to the root grid (hanging
    interfaces); `degrees(domain, leaves)` overrides the uniform degree;
    refine_depth > 0 builds the mesh by uniform refinement of one root (h-levels
    for multigrid)."""
    if refine_depth:
        domain = g.Domain((0.0,) * dim, (1.0,) * dim, (1,) * dim)
        g.uniform_refine(domain, refine_depth)
    else:
        domain = _refined_grid(dim, cells, refine)
    deg = g.UniformDegree(degree)
    if degrees is not None:
        deg = g.ByLeafDegrees(degrees)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=deg, continuity="dg"))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    pmax = int(np.max(space.degrees))
    problem.add_cell_accumulator(u, u, Diffusion())
    problem.add_interface_accumulator(u, u, SIPGInterface(pmax, eta))
    exact = exact or quadratic_exact
    for axis in range(dim):
        for side in (g.LOWER, g.UPPER):
            bf = g.BoundaryFace(axis, side)
            problem.add_boundary_accumulator(bf, u, u, SIPGBoundary(pmax, eta))
            if rhs:
                problem.add_boundary_accumulator(bf, u, SIPGBoundaryLoad(pmax, exact, eta))
    if rhs:
        problem.add_cell_accumulator(u, ConstantLoad(-6.0 if dim == 2 else -8.0) if source is None else source)
    return problem


class OneSidedAffine:
    """Nonsymmetric one-sided form of the reference's accumulator tests
    (test/accumulators.jl:70-87): a0 = m u + t d_n u, a_grad = s u n + k grad u."""

    def __init__(self, m, t, s, k):
        self.m, self.t, self.s, self.k = m, t, s, k

    def boundary_accumulate(self, q, trial, test_component):
        n = q.normal
        dn = g.normal_component(trial.gradient, n)
        return g.TestChannels(self.m * trial.value + self.t * dn,
                              tuple(self.s * trial.value * c + self.k * gc for c, gc in zip(n, trial.gradient)))


class InterfaceAffine:
    """Nonsymmetric interface form (test/accumulators.jl:105-113)."""

    def __init__(self, penalty, trial_normal, test_normal):
        self.penalty, self.trial_normal, self.test_normal = penalty, trial_normal, test_normal

    def interface_accumulate(self, q, trial, test_component):
        n = q.normal
        jmp = g.jump(trial.value)
        avg = g.average(g.normal_component(trial.gradient, n))
        tn = tuple(0.5 * self.test_normal * jmp * c for c in n)
        v = self.penalty * jmp + self.trial_normal * avg
        return g.TraceTestChannels(g.TestChannels(-v, tn), g.TestChannels(v, tn))


class GradientJumpPenalty:
    """Continuous interior penalty gamma h^2 [d_n u][d_n v] (acts on CG spaces too)."""

    def __init__(self, gamma=0.1):
        self.gamma = gamma

    def interface_accumulate(self, q, trial, test_component):
        n = q.normal
        h = q.cell_size[q.face_axis]
        jn = g.jump(g.normal_component(trial.gradient, n))
        c = self.gamma * h * h * jn
        return g.TraceTestChannels(g.TestChannels(0.0 * jn, tuple(-c * a for a in n)),
                                   g.TestChannels(0.0 * jn, tuple(c * a for a in n)))


# ---- nonlinear residual problems (examples/nonlinear_bratu/driver.jl) -------------------

class Bratu:
    """-Laplace u + lam exp(u) = f: residual (lam e^u - f, grad u), tangent
    (lam e^u du, grad du) (examples/nonlinear_bratu/driver.jl:38-52)."""

    def __init__(self, lam=1.0, source=None):
        self.lam, self.source = lam, source

    def cell_residual_accumulate(self, q, state, test_component):
        f = 0.0 if self.source is None else self.source(q.point)
        return g.TestChannels(self.lam * np.exp(state.value) - f, tuple(state.gradient))

    def cell_tangent_accumulate(self, q, state, increment, test_component):
        return g.TestChannels(self.lam * np.exp(state.value) * increment.value, tuple(increment.gradient))


def bratu_exact(amplitude):
    def exact(p):
        out = amplitude
        for x in p:
            out = out * np.sin(np.pi * x)
        return out
    return exact


def bratu_source(lam, amplitude, dim=2):
    exact = bratu_exact(amplitude)

    def source(p):
        e = exact(p)
        return dim * np.pi ** 2 * e + lam * np.exp(e)
    return source


def bratu_problem(root_counts=(4, 4), degree=3, lam=1.0, amplitude=0.2, refine=()):
    from gricoamd.nonlinear import ResidualProblem
    dim = len(root_counts)
    domain = g.Domain((0.0,) * dim, (1.0,) * dim, root_counts)
    last = None
    for cell, axis in refine:
        last = domain.refine(last if cell == -1 else cell, axis)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = ResidualProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Bratu(lam, bratu_source(lam, amplitude, dim)))
    for axis in range(dim):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return problem
