"""PDEGrid2D delta calls on the resident field (halhip_diffusion2d_delta_*)
against the same calls composed in FP64: the oracle's interface deltas for the
stencils (Diffusion2 uniform / Dirichlet edges / per-cell rates, Advection2
uniform / per-cell, Diffusion2DADI) and the reference's one-line delta rules
for Set / Add / Mul / Scale / SetAll / AddAll / MulAll / SetNonNegative
(include/Hal/Grids/PDEGrid2D.h:89-245), with MaxDelta, MaxDeltaScaled,
GetMax, GetMin and GetAvg, and Update's commit (:178-186).

Every stencil reads the committed field; Set overwrites the deltas at its
point in the call order.  Tolerance: FP32 arithmetic against FP64, 2e-6 of
the field's scale."""
import numpy as np
import pytest

import halhip
from helpers import RefGrid

pytestmark = pytest.mark.gpu


def edges(rng, X, Y):
    return (rng.random(Y), rng.random(Y), rng.random(X), rng.random(X))


@pytest.mark.parametrize("wrap", [(0, 0), (0, 1), (1, 1)])
def test_delta_calls_compose_like_pdegrid2d(ctx, oracle, wrap):
    X, Y = 40, 36
    rng = np.random.default_rng(17 + wrap[0] * 2 + wrap[1])
    u0 = rng.random(X * Y).astype(np.float32)
    p = halhip.Diffusion2D(ctx, X, Y)
    adi = halhip.Adi2D(ctx, X, Y)
    p.upload(u0)
    ref = RefGrid(oracle, u0, X, Y, wrap)
    bc = edges(rng, X, Y)
    vbc = tuple(0.2 * (e - 0.5) for e in edges(rng, X, Y))
    xv = 0.2 * (rng.random(X * Y) - 0.5)
    yv = 0.2 * (rng.random(X * Y) - 0.5)
    rx = 0.2 * rng.random(X * Y)
    ry = 0.2 * rng.random(X * Y)
    scale = 4.0

    def check(tag):
        got = p.deltas().astype(np.float64)
        np.testing.assert_allclose(got, ref.d, rtol=0, atol=2e-6 * scale, err_msg=tag)

    # tick 1: stencils with Dirichlet edges, point ops (Set overwrites), field ops
    p.delta_diffuse(0.2, *wrap, bc=bc)
    ref.diffuse(0.2, bc=bc)
    check("diffuse")
    p.delta_advect_cells(xv, yv, *wrap, bc=bc, vel_bc=vbc)
    ref.advect_cells(xv, yv, bc=bc, vbc=vbc)
    check("advect_cells")
    ops = [(halhip.DELTA_MUL, 5, 0.5), (halhip.DELTA_SET, 77, 3.0), (halhip.DELTA_ADD, 77, 1.0),
           (halhip.DELTA_SCALE, 300, 1.5), (halhip.DELTA_ADD, 5, -0.25)]
    p.delta_point_ops([o for o, _, _ in ops], [i for _, i, _ in ops], [v for _, _, v in ops])
    for o, i, v in ops:
        ref.point(o, i, v)
    check("point ops")
    p.delta_all(halhip.DELTA_ADD, 0.01)
    ref.d += 0.01
    p.delta_all(halhip.DELTA_MUL, -1e-3)
    ref.d += ref.u * -1e-3
    p.delta_adi(adi, 0.7, *wrap, bc=0.25)
    ref.adi(0.7, bc=0.25)
    check("adi")
    assert abs(p.delta_reduce(halhip.REDUCE_MAX_DELTA) - np.abs(ref.d).max()) < 2e-6 * scale
    scaled = np.abs(ref.d / (np.abs(ref.u) + 0.1)).max()
    assert abs(p.delta_reduce(halhip.REDUCE_MAX_DELTA_SCALED, 0.1) - scaled) < 1e-5 * scaled
    p.delta_all(halhip.DELTA_SET, -0.5)  # SetAll: every delta overwritten
    ref.d = -0.5 - ref.u
    p.delta_point_ops([halhip.DELTA_SET], [11], [2.0])
    ref.point(halhip.DELTA_SET, 11, 2.0)
    p.delta_diffuse(0.1, *wrap)
    ref.diffuse(0.1)
    p.delta_advect(0.2, -0.15, *wrap, bc=0.3)
    ref.advect(0.2, -0.15, bc=0.3)
    p.delta_all(halhip.DELTA_SET_NON_NEGATIVE)
    neg = ref.u + ref.d < 0
    ref.d[neg] = -ref.u[neg]
    check("tick 1 before Update")
    p.delta_update()
    ref.update()
    np.testing.assert_allclose(p.download().astype(np.float64), ref.u, rtol=0, atol=2e-6 * scale)
    assert np.abs(p.deltas()).max() == 0.0

    # tick 2: per-cell rates, SetAll from values, the field reductions
    p.delta_diffuse_rates(rx, ry, *wrap)
    ref.diffuse_rates(rx, ry)
    check("rates")
    vals = rng.random(X * Y)
    p.delta_set_values(vals)
    ref.d = vals - ref.u
    p.delta_adi(adi, 2.0, *wrap)
    ref.adi(2.0)
    p.delta_all(halhip.DELTA_SCALE, 0.9)
    ref.d += ref.u * (0.9 - 1)
    check("tick 2")
    u = ref.u
    assert abs(p.delta_reduce(halhip.REDUCE_FIELD_MAX) - u.max()) < 2e-6 * scale
    assert abs(p.delta_reduce(halhip.REDUCE_FIELD_MIN) - u.min()) < 2e-6 * scale
    assert abs(p.delta_reduce(halhip.REDUCE_FIELD_SUM) / u.size - u.mean()) < 2e-6 * scale
    p.delta_update()
    ref.update()
    np.testing.assert_allclose(p.download().astype(np.float64), ref.u, rtol=0, atol=4e-6 * scale)
    # a new committed field clears the pending deltas
    p.delta_all(halhip.DELTA_ADD, 1.0)
    p.upload(u0)
    assert np.abs(p.deltas()).max() == 0.0
    p.close()
    adi.close()


def test_delta_point_ops_reject_bad_input(ctx):
    p = halhip.Diffusion2D(ctx, 8, 8)
    with pytest.raises(halhip.InvalidArgument):
        p.delta_point_ops([halhip.DELTA_ADD], [64], [1.0])
    with pytest.raises(halhip.InvalidArgument):
        p.delta_point_ops([9], [0], [1.0])
    with pytest.raises(halhip.InvalidArgument):
        p.delta_all(99, 1.0)
    adi = halhip.Adi2D(ctx, 8, 9)
    with pytest.raises(halhip.InvalidArgument):
        p.delta_adi(adi, 1.0, False, False)
    p.close()
    adi.close()
