"""Learned per-cell block correction: MFMA network kernels against the MLP
oracle (forward and gradient, BF16 rounding emulated), cell patches against a
host gather, training, and the trained correction inside FGMRES against the
same method without it."""

import numpy as np
import pytest

import gricoamd as g
import mlp_oracle as MO
from gricoamd.learned import CellMLP, LearnedBlockCorrection, cell_weights, init_parameters, train_block_correction
from gricoamd.multigrid import GeometricMultigridPreconditioner

import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("widths,rows", [((64, 256, 256, 64), 300), ((27, 64, 48, 27), 77), ((125, 128, 128, 125), 1000)])
def test_forward_matches_oracle(gpu, widths, rows):
    rng = np.random.default_rng(0)
    p = init_parameters(widths, seed=5)
    p[-widths[-1]:] = rng.standard_normal(widths[-1]).astype(np.float32) * 0.1
    net = CellMLP(widths, parameters=p)
    x = rng.standard_normal((rows, widths[0])).astype(np.float32)
    y = net.forward(gpu.as_tensor(x)).cpu().numpy()
    ref, _ = MO.forward(p, widths, x, bf16=True)
    assert np.abs(y - ref).max() <= 1e-2 * np.abs(ref).max()
    # and close to the exact float64 network
    ref64, _ = MO.forward(p, widths, x, bf16=False)
    assert np.abs(y - ref64).max() <= 3e-2 * np.abs(ref64).max()


@pytest.mark.parametrize("widths,rows", [((64, 256, 256, 64), 512), ((27, 64, 48, 27), 130)])
def test_gradient_matches_oracle(gpu, widths, rows):
    rng = np.random.default_rng(1)
    p = init_parameters(widths, seed=7)
    net = CellMLP(widths, parameters=p)
    x = rng.standard_normal((rows, widths[0])).astype(np.float32)
    t = rng.standard_normal((rows, widths[-1])).astype(np.float32)
    loss, grad = net.gradient(gpu.as_tensor(x), gpu.as_tensor(t))
    rl, rg = MO.loss_and_gradient(p, widths, x, t, bf16=True)
    assert loss == pytest.approx(rl, rel=1e-3)
    # per-layer relative error in the Euclidean norm
    o = 0
    for k, n in zip(widths[:-1], widths[1:]):
        for size in (k * n, n):
            a, b = grad[o:o + size], rg[o:o + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7
            o += size


def test_training_fits_linear_map(gpu):
    """Adam on a noiseless linear target reduces the loss by orders of magnitude."""
    widths = (27, 64, 64, 27)
    rng = np.random.default_rng(2)
    A = rng.standard_normal((27, 27)).astype(np.float32) / np.sqrt(27)
    x = rng.standard_normal((8192, 27)).astype(np.float32)
    t = x @ A
    net = CellMLP(widths, seed=3)
    hist = net.train(gpu.as_tensor(x), gpu.as_tensor(t), batch=1024, steps=1500, learning_rate=3e-3)
    assert hist[-50:].mean() < 0.1 * hist[:10].mean()


def test_cell_patches(gpu):
    from gricoamd import _native, ops
    from gricoamd.learned import _declare
    plan = g.compile(P.lognormal_poisson(depth=1, degree=3))
    b = plan.buckets[0]
    n = plan.reduced_dof_count
    r = P.rng_vector(n, 1)
    e = P.rng_vector(n, 2)
    cw = cell_weights(plan)
    cells = np.array([0, 3, 5, 7, 7], dtype=np.int64)
    lib = _native.device()
    _declare(lib)
    tr, te = gpu.as_tensor(r, device="cuda"), gpu.as_tensor(e, device="cuda")
    tc = gpu.as_tensor(cells, device="cuda")
    X = gpu.empty((len(cells), b.n_modes), dtype=gpu.float32, device="cuda")
    T = gpu.empty_like(X)
    S = gpu.empty(len(cells), dtype=gpu.float32, device="cuda")
    tw = gpu.as_tensor(cw, device="cuda")
    from gricoamd import _devabi
    _devabi.check(lib.gmf_cell_patches(ops._ws(plan).ws, tr.data_ptr(), te.data_ptr(), tc.data_ptr(), len(cells),
                                       tw.data_ptr(), X.data_ptr(), T.data_ptr(), S.data_ptr()))
    red = plan.reduced_index
    for row, c in enumerate(cells):
        dofs = b.local_index[c]
        ridx = np.where(dofs >= 0, red[np.maximum(dofs, 0)], -1)
        loc_r = np.where(ridx >= 0, r[np.maximum(ridx, 0)], 0.0)
        loc_e = np.where(ridx >= 0, e[np.maximum(ridx, 0)], 0.0)
        s = np.linalg.norm(loc_r)
        assert S[row].item() == pytest.approx(s, rel=1e-6)
        assert np.allclose(X[row].cpu().numpy(), loc_r / s, atol=1e-6)
        assert np.allclose(T[row].cpu().numpy(), loc_e * cw[c] / s, rtol=1e-5, atol=1e-6)


def test_untrained_correction_keeps_fgmres_exact(gpu):
    """Any correction is admissible in FGMRES: the solution matches GMG-CG."""
    pr = P.lognormal_poisson(depth=2, degree=3, seed=4)
    net = CellMLP((64, 256, 256, 64), seed=0)
    st = g.solve(g.compile(pr), solver=g.FGMRESSolver(preconditioner=LearnedBlockCorrection(net)),
                 relative_tolerance=1e-12)
    ref = g.solve(pr, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner(smoother="chebyshev")),
                  relative_tolerance=1e-12)
    assert st.info.converged == 1
    # both stop at a 1e-12 relative residual of a problem whose coefficient
    # spans e^(+-3): the solutions agree to the conditioning times that
    assert np.abs(st.coefficients - ref.coefficients).max() <= 1e-7 * np.abs(ref.coefficients).max()


def test_trained_correction_reduces_iterations(gpu):
    """Trained in the energy norm on V-cycle residuals of GMG-CG solves, the
    correction leaves less energy error than it starts with and lowers the
    FGMRES iteration count on the problem's right-hand side, against FGMRES
    with the V-cycle alone (the same method without the network)."""
    pr = P.lognormal_poisson(depth=3, degree=3, seed=5)
    plan = g.compile(pr)
    net, report, M, V = train_block_correction(pr, plan=plan, pairs=200, steps=1500, seed=1)
    assert report["energy_ratio_last"] < report["energy_ratio_first"]
    assert report["energy_ratio_last"] < 1.0
    with_net = g.solve(plan, solver=g.FGMRESSolver(preconditioner=M), relative_tolerance=1e-8)
    without = g.solve(plan, solver=g.FGMRESSolver(preconditioner=V), relative_tolerance=1e-8)
    print("iterations with / without the network:", with_net.info.iterations, without.info.iterations, report)
    assert with_net.info.converged == 1 and without.info.converged == 1
    assert with_net.info.iterations < without.info.iterations


def test_energy_gradient_matches_oracle(gpu):
    """One step of the energy-norm training objective (gmf_learned_gradient:
    patches, network, per-cell rescaling, shared-dof averaging, q = A c, loss
    and backpropagation) against the oracle restatement with the device's BF16
    rounding, on a 4^3 p = 3 log-normal plan."""
    import ctypes

    import pyoracle
    from gricoamd import _devabi, _native
    from gricoamd.learned import _declare
    from gricoamd.solve import compile_preconditioner
    pr = P.lognormal_poisson(depth=2, degree=3, seed=8)
    plan = g.compile(pr)
    widths = (64, 256, 256, 64)
    rng = np.random.default_rng(3)
    p = init_parameters(widths, seed=9)
    p[-64:] = rng.standard_normal(64).astype(np.float32) * 0.05
    net = CellMLP(widths, parameters=p)
    J = compile_preconditioner(g.JacobiPreconditioner(), plan)
    M = LearnedBlockCorrection(net, inner=J).compile(plan)
    A = pyoracle.reduced_matrix(plan)
    r = rng.standard_normal(plan.reduced_dof_count)
    e = np.linalg.solve(A, r)
    energy = float(e @ A @ e)
    lib = _native.device()
    _declare(lib)
    loss = ctypes.c_double()
    grad = np.zeros(net.parameter_count, dtype=np.float32)
    rt = gpu.as_tensor(r, device="cuda")
    _devabi.check(lib.gmf_learned_gradient(M.handle, rt.data_ptr(), energy, ctypes.byref(loss),
                                           grad.ctypes.data_as(ctypes.c_void_p)))
    b = plan.buckets[0]
    dofs = np.asarray(b.local_index)
    ridx = np.where(dofs >= 0, plan.reduced_index[np.maximum(dofs, 0)], -1)
    cw = cell_weights(plan).astype(np.float64)
    rl, rg = MO.energy_loss_and_gradient(p, widths, A, r, energy, ridx, cw, bf16=True)
    assert loss.value == pytest.approx(rl, rel=1e-4)
    o = 0
    for k, n in zip(widths[:-1], widths[1:]):
        for size in (k * n, n):
            a, ref = grad[o:o + size], rg[o:o + size]
            assert np.linalg.norm(a - ref) <= 2e-2 * np.linalg.norm(ref) + 1e-9
            o += size
    # and the exact (float64) objective, to the BF16 network's accuracy
    l64, g64 = MO.energy_loss_and_gradient(p, widths, A, r, energy, ridx, cw, bf16=False)
    assert loss.value == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(grad - g64) <= 5e-2 * np.linalg.norm(g64)
