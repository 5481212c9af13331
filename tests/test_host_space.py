"""Host plan compiler: numbering goldens of the reference's own space tests
(test/space.jl:292-341) and structural invariants of the continuity
compile (src/continuity.jl)."""

import numpy as np
import pytest

import gricoamd as g


def terms(space, leaf, mode):
    return [(i, pytest.approx(c, abs=1e-14)) for i, c in space.mode_terms(leaf, mode)]


def test_single_cell_axis_degrees_numbering():
    """Compiled Leaf Regression: 6 dofs numbered in axis-1-fastest mode order."""
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.AxisDegrees((2, 1))))
    assert space.scalar_dof_count() == 6
    for dof, mode in enumerate([(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]):
        assert space.mode_terms(0, mode) == [(dof, 1.0)]


def test_hanging_face_golden():
    """Two roots, the second bisected along y: hanging trace modes tie to the
    coarse face with coefficients 0.5, 0.5, psi_2(0) = -sqrt(6)/4."""
    domain = g.Domain((0.0, 0.0), (2.0, 1.0), (2, 1))
    domain.refine(1, 1)
    space = g.HpSpace(domain, g.SpaceOptions(degree=g.UniformDegree(2)))
    assert space.scalar_dof_count() == 16
    want = [(1, 0.5), (4, 0.5), (7, -0.6123724356957945)]
    for leaf, mode in ((2, (0, 1)), (3, (0, 0))):
        got = space.mode_terms(leaf, mode)
        assert [i for i, _ in got] == [i for i, _ in want]
        assert np.allclose([c for _, c in got], [c for _, c in want], atol=1e-14)


def test_dg_numbering():
    domain = g.Domain((0.0,), (1.0,), (2,))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(3),
                                             continuity="dg"))
    assert space.scalar_dof_count() == 8
    assert space.scalar_dof_count() == sum(space.local_mode_count(l) for l in space.leaves)
    assert space.mode_terms(0, (1,)) == [(1, 1.0)]
    assert space.mode_terms(1, (0,)) == [(4, 1.0)]
    p0 = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(0), continuity="dg"))
    assert p0.scalar_dof_count() == 2
    assert p0.cell_quadrature_shape(0) == (1,)
    assert p0.local_mode_count(0) == 1
    assert p0.mode_terms(0, (0,)) == [(0, 1.0)]
    assert p0.mode_terms(1, (0,)) == [(1, 1.0)]


def test_quadrature_shape_default():
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.AxisDegrees((2, 1))))
    assert space.cell_quadrature_shape(0) == (3, 2)
    assert space.is_mode_active(0, (2, 1))


@pytest.mark.parametrize("dim,cells,p", [(2, 4, 3), (3, 3, 2), (3, 2, 4)])
def test_uniform_cg_dof_counts(dim, cells, p):
    domain = g.Domain((0.0,) * dim, (1.0,) * dim, (cells,) * dim)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(p)))
    assert space.scalar_dof_count() == (cells * p + 1) ** dim
    # every mode of a conforming uniform mesh is a single unit term
    assert np.all(np.diff(space.term_offsets) == 1)
    assert np.all(space.term_coefficients == 1.0)


def test_uniform_refine_matches_roots():
    """Six uniform refinements of one root give the same dof count as 64^3 roots
    would; the leaves are all at level 6 (benchmark/hpmg_scaling.jl:122-170)."""
    domain = g.Domain((0.0,) * 3, (1.0,) * 3, (1, 1, 1))
    g.uniform_refine(domain, 2)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    assert space.active_leaf_count() == 64
    assert space.scalar_dof_count() == 9 ** 3


def test_trunk_mode_counts():
    """TrunkBasis keeps the vertex modes and modes whose superlinear degree sum
    stays within p (src/basis.jl)."""
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    for p, n in ((1, 4), (2, 8), (3, 12), (4, 17)):
        space = g.HpSpace(domain, g.SpaceOptions(basis=g.TrunkBasis(), degree=g.UniformDegree(p)))
        assert space.local_mode_count(0) == n


def test_hanging_trace_agreement():
    """A continuous function from the coarse side equals the fine side's trace
    on the hanging face: for every coarse vector, the fine cells' face values
    interpolate the coarse cell's face values (test/space.jl:395-401)."""
    from gricoamd.quadrature import fe_basis_values_and_derivatives
    domain = g.Domain((0.0, 0.0), (2.0, 1.0), (2, 1))
    domain.refine(1, 1)
    p = 3
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(p)))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(space.scalar_dof_count())

    def value(leaf, xi, eta):
        vx, _ = fe_basis_values_and_derivatives([xi], p)
        vy, _ = fe_basis_values_and_derivatives([eta], p)
        out = 0.0
        for i in range(p + 1):
            for j in range(p + 1):
                for dof, c in space.mode_terms(leaf, (i, j)):
                    out += c * x[dof] * vx[0][i] * vy[0][j]
        return out

    for y in np.linspace(0.0, 1.0, 9):
        coarse = value(0, 1.0, 2 * y - 1)
        if y <= 0.5:
            fine = value(2, -1.0, 4 * y - 1)
        else:
            fine = value(3, -1.0, 4 * (y - 0.5) - 1)
        assert fine == pytest.approx(coarse, abs=1e-12)
