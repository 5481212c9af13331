"""Finite-cell geometry on the host: leaf classification, cut-cell candidate
rules (src/regions.jl:304-329, 477-517, 736-817) and the sphere surface rule,
checked against the closed-form volume and area of a ball."""

import numpy as np
import pytest

import gricoamd as g
from gricoamd.embedded import (CUT, INSIDE, OUTSIDE, ImplicitRegion, add_cut_cell_quadratures, classify_leaves,
                               cut_cell_rules, physical_space, sphere_surface_quadratures)


def ball(r=1.0):
    return ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - r * r, subdivision_depth=3)


def unit_ball_domain(n=12):
    return g.Domain((-1.2, -1.2, -1.2), (2.4, 2.4, 2.4), (n, n, n))


def test_classification_counts():
    dom = unit_ball_domain(12)
    cls = classify_leaves(dom, ball(), dom.active_leaves())
    assert set(np.unique(cls)) <= {INSIDE, OUTSIDE, CUT}
    # centre cell inside, corner cell outside
    lo = dom.cell_lowers(dom.active_leaves())
    centre = np.argmin(np.linalg.norm(lo + 0.1, axis=1))
    corner = np.argmin(np.linalg.norm(lo + 1.2, axis=1))
    assert cls[centre] == INSIDE and cls[corner] == OUTSIDE
    assert (cls == CUT).sum() > 0


@pytest.mark.parametrize("depth,tol", [(1, 3e-2), (3, 6e-3)])
def test_volume_converges(depth, tol):
    dom = unit_ball_domain(12)
    region = ImplicitRegion(ball().classifier, subdivision_depth=depth)
    leaves = dom.active_leaves()
    cls = classify_leaves(dom, region, leaves)
    h = 2.4 / 12
    vol = (cls == INSIDE).sum() * h ** 3
    rules = cut_cell_rules(dom, region, leaves[cls == CUT], (3, 3, 3))
    vol += sum(r.weights.sum() for r in rules.values()) * (h / 2) ** 3
    assert abs(vol - 4.0 / 3.0 * np.pi) <= tol * 4.0 / 3.0 * np.pi
    for r in rules.values():
        assert np.all(np.abs(r.points) <= 1.0) and np.all(r.weights > 0)


def test_sphere_surface_area_and_normals():
    dom = unit_ball_domain(12)
    space, cut = physical_space(dom, ball(), g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    for sq in sphere_surface_quadratures(space, (0, 0, 0), 1.0, n_theta=32, n_phi=64):
        pr.add_surface_quadrature(sq)
    plan = g.compile(pr)
    area = sum(b.weights.sum() for b in plan.surface_buckets)
    assert area == pytest.approx(4 * np.pi, rel=1e-12)
    for b in plan.surface_buckets:
        x = np.stack(b.phys_points, axis=1)
        assert np.allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-12)
        assert np.allclose(b.normals, x, atol=1e-12)


def test_cut_rules_attach():
    dom = unit_ball_domain(8)
    region = ImplicitRegion(ball().classifier, subdivision_depth=2)
    space, cut = physical_space(dom, region, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    n = add_cut_cell_quadratures(pr, region, cut)
    assert n > 0 and len(pr.cell_quadratures) == len(cut)
    plan = g.compile(pr)
    assert sum(b.cell_count for b in plan.cloud_buckets) == len(cut)


# ---- physical domains, sampled surfaces, point location --------------------------

def test_locate_points_on_mixed_leaf_sizes():
    from gricoamd.embedded import locate_points
    dom = g.Domain((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (2, 1, 1))
    first = dom.refine(1, 0)
    dom.refine(first, 1)
    leaves = np.asarray(dom.active_leaves())
    rng = np.random.default_rng(3)
    x = rng.uniform([0, 0, 0], [2, 1, 1], size=(400, 3))
    owner = locate_points(dom, leaves, x)
    assert np.all(owner >= 0)
    lo = dom.cell_lowers(leaves)[owner]
    hi = lo + dom.cell_sizes(leaves)[owner]
    assert np.all((x >= lo - 1e-12) & (x <= hi + 1e-12))
    # a brute-force check of ownership
    for p_, o in zip(x[:50], owner[:50]):
        inside = np.all((p_ >= dom.cell_lowers(leaves)) & (p_ < dom.cell_lowers(leaves) + dom.cell_sizes(leaves)), axis=1)
        assert inside[o]
    # points outside the leaf set are not located; the domain's upper faces are
    out = locate_points(dom, leaves[:1], np.array([[1.5, 0.5, 0.5], [0.2, 0.2, 0.2]]))
    assert out[0] == -1 and out[1] == 0
    assert locate_points(dom, leaves, np.array([[2.0, 1.0, 1.0]]))[0] >= 0


def test_sampled_surface_pieces_on_mixed_leaf_sizes():
    # every sample lands inside its piece's leaf (nothing is clipped onto a
    # face), and the pieces keep every sample once
    from gricoamd.embedded import sphere_surface
    dom = g.Domain((-1.2,) * 3, (2.4,) * 3, (3, 3, 3))
    for axis in range(3):
        dom.refine(int(dom.active_leaves()[0]), axis)
    dom.refine(int(dom.active_leaves()[13]), 0)
    sp = g.HpSpace(dom, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(1)))
    surface = sphere_surface((0.05, -0.1, 0.0), 1.0, n_theta=12, n_phi=24)
    pieces = surface.surface_quadratures(sp)
    leaves = list(np.asarray(sp.leaves))
    got = []
    for piece in pieces:
        li = leaves.index(piece.leaf)
        lo = dom.cell_lowers([piece.leaf])[0]
        half = 0.5 * dom.cell_sizes([piece.leaf])[0]
        got.append(lo + half * (piece.quadrature.points + 1.0))
    got = np.concatenate(got)
    assert len(got) == len(surface.points)
    key = lambda a: np.lexsort(a.T[::-1])
    assert np.allclose(got[key(got)], surface.points[key(surface.points)], atol=1e-12)


def test_finite_cell_extension_rules_and_checks():
    from gricoamd.embedded import (FiniteCellExtension, ImplicitRegion, PhysicalDomain, PhysicalMeasure, CUT,
                                   _tensor_rule)
    with pytest.raises(ValueError):
        FiniteCellExtension(1.5)
    with pytest.raises(ValueError):
        FiniteCellExtension(float("nan"))
    region = ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=1)
    dom = g.Domain((-1.2,) * 3, (2.4,) * 3, (3, 3, 3))
    opts = g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(1))
    phys = PhysicalDomain(dom, region, PhysicalMeasure())
    ext = PhysicalDomain(dom, region, FiniteCellExtension(0.25))
    full = PhysicalDomain(dom, region, FiniteCellExtension(1.0))
    sp = g.HpSpace(phys, opts)
    # outside leaves are dropped, and the space's leaves agree across measures
    assert np.all(phys.classify(sp.leaves) != 1)
    rp, re_ = phys.assembly_cell_quadratures(sp), ext.assembly_cell_quadratures(g.HpSpace(ext, opts))
    assert set(rp) == set(re_) and len(rp) > 0
    assert not full.assembly_cell_quadratures(g.HpSpace(full, opts))  # alpha = 1: the tensor rule everywhere
    fp, fw = _tensor_rule((2, 2, 2))
    for leaf, r in rp.items():
        e = re_[leaf]
        npnt = r.point_count()
        assert np.allclose(e.weights[:npnt], 0.75 * r.weights) and np.allclose(e.weights[npnt:], 0.25 * fw)
        assert np.allclose(e.points[npnt:], fp)
    # rules are cached per leaf and shape
    leaf = next(iter(rp))
    assert phys.physical_rules([leaf], (2, 2, 2))[leaf] is phys.physical_rules([leaf], (2, 2, 2))[leaf]


def test_conflict_coloring_is_valid():
    import ctypes
    from gricoamd import _native
    lib = _native.host()
    lib.gh_conflict_coloring.argtypes = [ctypes.c_int64, _native.P_i64, _native.P_i64, ctypes.c_int64,
                                         _native.P_i64, _native.P_i64, ctypes.c_int64, _native.P_i32,
                                         ctypes.POINTER(ctypes.c_int32)]
    rng = np.random.default_rng(5)
    ndof = 300
    items = [np.unique(rng.integers(0, ndof, size=int(rng.integers(1, 9)))) for _ in range(120)]
    blocks = [np.unique(rng.integers(0, ndof, size=int(rng.integers(1, 6)))) for _ in range(80)]
    iptr = np.concatenate([[0], np.cumsum([len(i) for i in items])]).astype(np.int64)
    idofs = np.concatenate(items).astype(np.int64)
    bptr = np.concatenate([[0], np.cumsum([len(b) for b in blocks])]).astype(np.int64)
    bdofs = np.concatenate(blocks).astype(np.int64)
    colors = np.zeros(len(blocks), dtype=np.int32)
    count = ctypes.c_int32()
    _native.check_host(lib.gh_conflict_coloring(len(blocks), _native.ptr(bptr, _native.c_int64),
                                                _native.ptr(bdofs, _native.c_int64), len(items),
                                                _native.ptr(iptr, _native.c_int64), _native.ptr(idofs, _native.c_int64),
                                                ndof, _native.ptr(colors, _native.c_int32), ctypes.byref(count)))
    assert colors.max() + 1 == count.value
    for it in items:
        touched = [b for b, bl in enumerate(blocks) if np.intersect1d(it, bl).size]
        assert len({int(colors[b]) for b in touched}) == len(touched)
    # greedy: every block's color is the smallest one its conflicting predecessors leave free
    for b in range(len(blocks)):
        conflicting = {int(colors[o]) for o in range(b) for it in items
                       if np.intersect1d(it, blocks[b]).size and np.intersect1d(it, blocks[o]).size}
        assert colors[b] == min(set(range(len(blocks) + 1)) - conflicting)


def test_point_cloud_load_layout_reproduces_host_rhs():
    # The loads handed to gmf_rhs_faces (per point-cloud batch, item, point:
    # W v, W g_1..g_D) projected with the host's own bucket projection and
    # scattered give plan.rhs_physical(): the layout the device kernel reads.
    import problems as P
    pr = P.fcm_ball(n=4, degree=2, depth=2, n_theta=12, n_phi=24, alpha=1e-3)
    pr.add_surface_accumulator(pr.cell_operators[0][0], P.NitscheData(90.0))
    plan = g.compile(pr)
    assert plan.device_rhs and plan.cloud_buckets and plan.surface_buckets
    D = plan.dim
    b = np.zeros(plan.full_dof_count)
    loads, ch = plan.lowered_loads()
    off = 0
    for bk in plan.buckets:
        n = bk.cell_count * bk.n_points * ch
        arr = loads[off:off + n].reshape(bk.cell_count, bk.n_points, ch)
        off += n
        local = bk.project(arr[..., 0])
        for a in range(ch - 1):
            local = local + bk.project(arr[..., 1 + a], derivative_axis=a)
        b += np.bincount(bk.term_dof, weights=bk.term_coef * local[bk.term_cell, bk.term_box],
                         minlength=plan.full_dof_count)
    fl = plan.face_loads()
    off = 0
    for bk in plan.all_buckets:
        if not getattr(bk, "cloud", False):
            continue
        n = bk.n_points_total * (D + 1)
        arr = fl[off:off + n].reshape(bk.n_points_total, D + 1)
        off += n
        local = bk.project(arr[:, 0])
        for a in range(D):
            local = local + bk.project(arr[:, 1 + a], derivative_axis=a)
        b += np.bincount(bk.term_dof, weights=bk.term_coef * local[bk.term_cell, bk.term_box],
                         minlength=plan.full_dof_count)
    assert off == fl.size
    ref = plan.rhs_physical()
    assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()
    # the surface data terms contribute (value and gradient channels)
    assert np.abs(fl.reshape(-1, D + 1)[:, 1:]).max() > 0


def test_l2_layout_reproduces_host_error():
    # gmf_l2_error's per-point weights and exact values (tensor cells, then
    # one-sided point-cloud batches, zero weights off the cells) summed with
    # the host's own field values give the host relative L2 error.
    import problems as P
    from gricoamd.verification import _field_values, relative_l2_error
    pr = P.fcm_ball(n=4, degree=2, depth=2, n_theta=12, n_phi=24, alpha=1e-3)
    plan = g.compile(pr)
    rng = np.random.default_rng(2)
    st = g.State(plan, rng.standard_normal(plan.full_dof_count))
    w = plan.point_weights()
    e = plan.point_values(P.ball_exact)
    off, num, den = 0, 0.0, 0.0
    for bk, counted in plan._l2_layout():
        n = bk.n_points_total if getattr(bk, "cloud", False) else bk.cell_count * bk.n_points
        uh = _field_values(plan, st.coefficients, bk).reshape(-1) if counted else np.zeros(n)
        ww, ee = w[off:off + n], e[off:off + n]
        num += float(np.sum(ww * (uh - ee) ** 2))
        den += float(np.sum(ww * ee ** 2))
        off += n
    assert off == w.size == e.size
    assert any(not c for _, c in plan._l2_layout())  # the surface pieces take zero weights
    ref = relative_l2_error(st, P.ball_exact)
    assert abs(np.sqrt(num / den) - ref) <= 1e-12 * ref


def test_brick_grouping_of_ball_tensor_cells():
    # host grouping for the device brick kernel against a dictionary built
    # cell by cell: (level, coordinate // 2) per axis names the brick, the
    # coordinate parities the position; a complete Morton-ordered mesh needs none
    import problems as P
    from gricoamd.plan import brick_grouping
    pr = P.fcm_ball(n=8, degree=1, depth=1, n_theta=12, n_phi=24)
    plan = g.compile(pr)
    b = next(b for b in plan.buckets if not getattr(b, "cloud", False))
    grid = pr.field.space.domain.grid
    lev, crd = grid.levels()[b.cells], grid.coords()[b.cells]
    want = {}
    for i in range(b.cell_count):
        key = tuple(lev[i]) + tuple(crd[i] // 2)
        want.setdefault(key, [-1] * 8)[int(crd[i][0] % 2 + 2 * (crd[i][1] % 2) + 4 * (crd[i][2] % 2))] = i
    got = b.brick_cells
    assert got is not None and got.shape == (len(want), 8)
    assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, want.values()))
    firsts = [min(c for c in row if c >= 0) for row in got.tolist()]
    assert firsts == sorted(firsts)  # bricks in the order of their first cell
    assert (got < 0).any() and sorted(got[got >= 0].tolist()) == list(range(b.cell_count))
    pr2 = P.poisson(3, 4, 1, rhs=False)
    full = g.compile(pr2)
    assert full.buckets[0].brick_cells is None
    # the same cells in reverse order need an explicit grouping
    rev = brick_grouping(pr2.field.space.domain.grid, full.buckets[0].cells[::-1])
    assert rev is not None and rev.shape == (8, 8) and (rev >= 0).all()
