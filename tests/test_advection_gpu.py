"""Upwind advection on the GPU (PDEGrid2D::Advection, uniform and per-cell
velocity, through the C ABI) against the compiled reference's golden
vectors and the FP64 oracle's model ticks (Diffusion + Advection + MulAll +
Add + Update).

Tolerances: the interface computes in FP32 from FP64 inputs, so single calls
are checked to 1e-6 absolute (values are O(1)); resident ticks to a relative
L2 of 1e-6 against FP64.
"""
import numpy as np
import pytest

import halhip
from helpers import advection_cases, rel_l2

pytestmark = pytest.mark.gpu


def test_interface_matches_golden(ctx, golden):
    plans = {}
    for c in advection_cases(golden):
        key = (c["X"], c["Y"])
        if key not in plans:
            plans[key] = halhip.Diffusion2D(ctx, *key)
        p = plans[key]
        got = p.advect(c["field"], c["deltas_in"].copy(), *c["vel"], c["wrapX"], c["wrapY"], bc=c["bc"])
        np.testing.assert_allclose(got, c["uniform_out"], rtol=0, atol=1e-6, err_msg=str(c["k"]))
        got = p.advect_cells(c["field"], c["deltas_in"].copy(), c["xv"], c["yv"], c["wrapX"], c["wrapY"],
                             bc=c["bc"], vel_bc=c["vbc"])
        np.testing.assert_allclose(got, c["cells_out"], rtol=0, atol=1e-6, err_msg=str(c["k"]))


def test_model_ticks_golden(ctx, golden):
    f = golden["ticks_model_field"]
    p = halhip.Diffusion2D(ctx, 20, 18)
    p.upload(f)
    p.set_sources([5, 200], [1.0, 0.5])
    p.set_advection(0.2, -0.1)
    p.run_steps(15, 0.15, 2e-3, True, False)
    assert rel_l2(p.download(), golden["ticks_model_out"]) < 1e-6
    p.upload(f)
    p.set_sources([7], [0.3])
    p.set_advection(-0.05, 0.25)
    p.run_steps(15, 0.15, 0.0, False, False, bc=0.4)
    assert rel_l2(p.download(), golden["ticks_model_bc_out"]) < 1e-6


# several ticks per launch (yDim <= 1024) and the single-tick kernel (longer
# rows); rate 0.1 with |xVel| + |yVel| <= 0.5 keeps the explicit scheme
# monotone (centre weight 1 - 4r - |xVel| - |yVel| >= 0), so FP32 rounding
# is not amplified over the ticks
MODEL_CASES = [
    (40, 64, 1, 1, None, 0.25, -0.2),
    (33, 100, 0, 0, 0.3, -0.1, 0.3),
    (17, 1024, 0, 1, None, 0.0, -0.45),
    (9, 1500, 1, 0, 0.5, 0.3, 0.1),
    (70, 31, 0, 0, "edges", 0.2, 0.2),
]


@pytest.mark.parametrize("X,Y,wx,wy,bc,vx,vy", MODEL_CASES)
def test_resident_model_ticks_match_oracle(ctx, oracle, X, Y, wx, wy, bc, vx, vy):
    rng = np.random.default_rng(X + 7 * Y)
    f = rng.random(X * Y).astype(np.float32)
    if bc == "edges":
        bc = tuple(rng.random(n) for n in (Y, Y, X, X))
    idx = rng.integers(0, X * Y, 4)
    val = rng.random(4).astype(np.float32)
    p = halhip.Diffusion2D(ctx, X, Y)
    p.upload(f)
    p.set_sources(idx, val)
    p.set_advection(vx, vy)
    p.run_steps(45, 0.1, 1e-3, wx, wy, bc=bc)
    ref = oracle.model_ticks(f.astype(np.float64), 45, 0.1, vx, vy, 1e-3, X, Y, wx, wy, bc, idx,
                             val.astype(np.float64))
    assert rel_l2(p.download(), ref) < 1e-6


def test_advection_errors(ctx):
    p = halhip.Diffusion2D(ctx, 4, 5)
    with pytest.raises(halhip.InvalidArgument):
        p.set_advection(0.3, -0.3)
    with pytest.raises(halhip.InvalidArgument):
        p.advect(np.zeros(20), np.zeros(19), 0.1, 0.1, False, False)
    with pytest.raises(halhip.InvalidArgument):
        p.advect_cells(np.zeros(20), np.zeros(20), np.zeros(20), np.zeros(20), False, False, bc=None, vel_bc=0.1)
    # setting the velocity back to zero restores plain diffusion
    f = np.random.default_rng(1).random(20).astype(np.float32)
    p.set_advection(0.2, 0.1)
    p.set_advection(0.0, 0.0)
    p.upload(f)
    p.run_steps(3, 0.2, 0.0, True, True)
    a = p.download()
    q = halhip.Diffusion2D(ctx, 4, 5)
    q.upload(f)
    q.run_steps(3, 0.2, 0.0, True, True)
    np.testing.assert_array_equal(a, q.download())
