"""Adjoint ADI with periodic and Dirichlet boundaries on the GPU: the forward
model against the FP64 oracle and the plain uniform-rate ADI, gradients
against the FP64 CPU adjoint and FP64 central finite differences, for every
boundary kind (wrapped axes; Dirichlet constant or edge arrays on the others).

Bars: forward rel L2 1e-5; the gradient within 1e-4 max|g| of FP64 central
differences and of the FP64 adjoint cell by cell, and within 2e-4 rel L2 of
the FP64 adjoint (the fully periodic grid's gradient cancels most: 1.0e-4
measured there, 8e-5 zero flux, ~1e-6 with Dirichlet edges); dJ/dk 1e-4."""
import numpy as np
import pytest

import halhip
from helpers import rel_l2

pytestmark = pytest.mark.gpu

# (wrapX, wrapY, bc): bc None (zero flux), a constant, or "edges"
BOUNDS = [(1, 1, None), (1, 0, None), (0, 1, None), (0, 0, 0.7), (1, 0, 0.4), (0, 1, -0.3),
          (0, 0, "edges"), (1, 0, "edges")]


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def make_bc(bc, X, Y, seed=5):
    """(bc for the device, the same ghost values as the FP32 device holds them
    for the oracle)."""
    if isinstance(bc, str):
        rng = np.random.default_rng(seed)
        e = (rng.random(Y), rng.random(Y), rng.random(X), rng.random(X))
        return e, tuple(f32(a) for a in e)
    return bc, (None if bc is None else float(np.float32(bc)))


def problem(X, Y, seed):
    rng = np.random.default_rng(seed)
    return 0.5 + rng.random(X * Y), rng.random(X * Y).astype(np.float32)


# Every per-thread width of both line solves (rows: 1..16 elements per thread;
# columns: 1..64 rows per lane), two-cell and one-cell wrapped axes.
@pytest.mark.parametrize("X,Y", [(24, 20), (130, 70), (300, 520), (1030, 18), (2052, 8), (2, 66), (1, 33),
                                 (40, 1), (520, 2100)])
def test_forward_every_boundary_matches_oracle(ctx, oracle, X, Y):
    D, u0 = problem(X, Y, X + Y)
    p = halhip.Adjoint2D(ctx, X, Y)
    for wx, wy, bc in BOUNDS:
        dev_bc, ora_bc = make_bc(bc, X, Y)
        p.set_problem(D, 5e-3, 12, 5, wx, wy, dev_bc)
        got = p.forward(u0)
        ref = oracle.adi_cell_forward(u0.astype(np.float64), f32(D), float(np.float32(5e-3)), 12, X, Y, wx, wy,
                                      ora_bc)
        err = rel_l2(got, ref)
        assert err < 1e-5, (wx, wy, bc, err)
    p.close()


@pytest.mark.parametrize("wx,wy,bc", [(1, 1, None), (1, 0, None), (0, 1, None), (0, 0, 0.7), (1, 0, 0.4),
                                      (0, 0, "edges")])
def test_uniform_diffusivity_matches_plain_adi_every_boundary(ctx, wx, wy, bc):
    X, Y = 50, 64
    u0 = np.random.default_rng(4).random(X * Y).astype(np.float32)
    dev_bc, _ = make_bc(bc, X, Y)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(np.full(X * Y, 2.0), 1e-3, 10, 10, wx, wy, dev_bc)
    plain = halhip.Adi2D(ctx, X, Y)
    plain.upload(u0)
    plain.run_steps(10, 2.0, 1e-3, wx, wy, bc=dev_bc)
    assert rel_l2(p.forward(u0), plain.download()) < 1e-6
    p.close()
    plain.close()


def make_target(oracle, u0, D, k, ticks, X, Y, wx, wy, bc, seed):
    # perturbed D and twice the decay: away from the misfit minimum in k
    rng = np.random.default_rng(seed)
    return f32(oracle.adi_cell_forward(u0.astype(np.float64), D * (1 + 0.1 * rng.random(X * Y)), 2 * k, ticks, X,
                                       Y, wx, wy, bc))


@pytest.mark.parametrize("wx,wy,bc", BOUNDS)
def test_gradient_every_boundary_matches_fp64_adjoint(ctx, oracle, wx, wy, bc):
    X, Y, ticks, k = 40, 36, 40, 5e-3
    D, u0 = problem(X, Y, 3)
    dev_bc, ora_bc = make_bc(bc, X, Y)
    target = make_target(oracle, u0, D, k, ticks, X, Y, wx, wy, ora_bc, 9)
    J, gD, gk = oracle.adi_cell_gradient(u0.astype(np.float64), f32(D), float(np.float32(k)), ticks, X, Y, target,
                                         wx, wy, ora_bc)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, k, ticks, 6, wx, wy, dev_bc)
    Jg, gDg, gkg = p.gradient(u0, target.astype(np.float32))
    p.close()
    err = rel_l2(gDg, gD)
    worst = np.abs(gDg - gD).max() / np.abs(gD).max()
    print(f"boundaries {wx}{wy} {bc}: dJ/dD rel L2 {err:.3e}, max |dg| / max|g| {worst:.3e}, "
          f"dJ/dk {abs(gkg - gk) / abs(gk):.3e}")
    assert abs(Jg - J) < 1e-4 * abs(J)
    assert err < 2e-4 and worst < 1e-4, (err, worst)
    assert abs(gkg - gk) < 1e-4 * abs(gk)


@pytest.mark.parametrize("wx,wy,bc", [(1, 1, None), (0, 0, "edges"), (1, 0, 0.4)])
def test_gradient_every_boundary_vs_central_differences(ctx, oracle, wx, wy, bc):
    # FP64 central differences (h = 1e-6) of the FP64 model at every corner,
    # cells of every edge and seeded interior cells.
    X, Y, ticks, k = 32, 28, 30, 5e-3
    D, u0 = problem(X, Y, 5)
    dev_bc, ora_bc = make_bc(bc, X, Y)
    target = make_target(oracle, u0, D, k, ticks, X, Y, wx, wy, ora_bc, 6)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, k, ticks, 10, wx, wy, dev_bc)
    _, gD, gk = p.gradient(u0, target.astype(np.float32))
    p.close()
    u64 = u0.astype(np.float64)
    Dd = f32(D)
    kd = float(np.float32(k))

    def loss(Dv, kv):
        return 0.5 * np.sum((oracle.adi_cell_forward(u64, Dv, kv, ticks, X, Y, wx, wy, ora_bc) - target) ** 2)

    h = 1e-6
    scale = np.abs(gD).max()
    rng = np.random.default_rng(7)
    cells = [0, Y - 1, (X - 1) * Y, X * Y - 1, 5, 7 * Y, 9 * Y + Y - 1, (X - 1) * Y + 11]
    cells += list(rng.choice(X * Y, 24, replace=False))
    worst = 0.0
    for c in cells:
        Dp, Dm = Dd.copy(), Dd.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (loss(Dp, kd) - loss(Dm, kd)) / (2 * h)
        worst = max(worst, abs(gD[c] - fd) / scale)
        assert abs(gD[c] - fd) < 1e-4 * scale, (c, gD[c], fd)
    fdk = (loss(Dd, kd + h) - loss(Dd, kd - h)) / (2 * h)
    print(f"boundaries {wx}{wy} {bc}: max |g - fd| / max|g| = {worst:.3e}, decay {abs(gk - fdk) / abs(fdk):.3e}")
    assert abs(gk - fdk) < 1e-4 * abs(fdk)


def test_wrapped_one_cell_axis_and_edge_array_lengths(ctx):
    # a wrapped axis of one cell has no faces: the same as zero flux there
    X, Y = 1, 40
    D, u0 = problem(X, Y, 8)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, 1e-3, 5, 5, 1, 0, None)
    a = p.forward(u0)
    p.set_problem(D, 1e-3, 5, 5, 0, 0, None)
    b = p.forward(u0)
    np.testing.assert_array_equal(a, b)
    p.close()
    # edge arrays of the wrong lengths, diffusivities and decays that FP32
    # cannot hold, negative or NaN diffusivities and zero ticks are refused,
    # and the problem set before stands
    X, Y = 6, 5
    D, u0 = problem(X, Y, 9)
    q = halhip.Adjoint2D(ctx, X, Y)
    q.set_problem(D, 1e-3, 5, 5, 1, 0, 0.25)
    before = q.forward(u0)
    bad_edges = (np.zeros(X), np.zeros(Y), np.zeros(X), np.zeros(X))   # x-edges hold yDim values
    cell = np.arange(X * Y)
    for args in ((D, 1e-3, 5, 5, 0, 0, bad_edges), (D, 1e-3, 5, 5, 0, 0, bad_edges[:3]),
                 (np.where(cell == 7, 1e40, D), 1e-3, 5, 5, 1, 0, None),
                 (np.where(cell == 3, -1.0, D), 1e-3, 5, 5, 1, 0, None),
                 (np.where(cell == 0, np.nan, D), 1e-3, 5, 5, 0, 1, None),
                 (D, 1e300, 5, 5, 1, 0, None), (D, 1e-3, 0, 5, 0, 1, None), (D[:-1], 1e-3, 5, 5, 1, 0, None)):
        with pytest.raises(halhip.InvalidArgument):
            q.set_problem(*args)
    np.testing.assert_array_equal(q.forward(u0), before)
    q.close()
