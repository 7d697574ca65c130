"""Agent binning and forces on the GPU vs the FP64 oracle, the reference's
golden vectors and brute force (IAgentForceBackend2D, AgentForceKernel2D).

Bins (offsets, bucket order) must be bit-identical to the reference's for the
same FP32 positions; forces and positions are FP32 results compared with
tolerances stated per test.
"""
import numpy as np
import pytest

import halhip
import workloads
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def test_golden_scene_through_interface(ctx, golden, oracle):
    # tests/IAgentForceBackend2DTests.cpp:51-121 (GPU tolerance 1e-3 there)
    pos, rad = golden["agents_pos"], golden["agents_rad"]
    plan = halhip.Agents2D(ctx, 16, 16, 60)
    for wrap, key in (((1, 1), "agents_wrap_out"), ((0, 0), "agents_nowrap_out")):
        got = plan.accumulate_forces(pos, rad, 3.0, 0.4, *wrap)
        np.testing.assert_allclose(got, golden[key], rtol=0, atol=2e-5)
        # Same FP32-rounded inputs through the oracle: the only difference left
        # is FP32 arithmetic.
        ref = oracle.agent_forces(f32(pos), f32(rad), 3.0, 0.4, 16, 16, *wrap)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5)


def test_bins_bit_exact_and_forces(ctx, golden, oracle):
    pos = golden["agents2_pos"]
    rad = np.full(500, 0.5)
    plan = halhip.Agents2D(ctx, 40, 24, 1000)
    got = plan.accumulate_forces(pos, rad, 1.0, 1.0, 0, 1)
    np.testing.assert_allclose(got, golden["agents2_out"], rtol=0, atol=2e-5)
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(f32(pos), 40, 24)
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)


def test_crowded_cells_order(ctx, oracle):
    # Many agents per cell exercise the per-cell ordering pass.
    rng = np.random.default_rng(7)
    n = 20000
    pos = (rng.random(2 * n) * 6.0 - 1.0).astype(np.float32)       # 6x6 grid, some clamped
    plan = halhip.Agents2D(ctx, 5, 4, n)
    plan.upload(pos, np.full(n, 0.1, np.float32))
    plan.run_ticks(0, 0.05, 1.0, 0.0, 0, 0)
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(pos.astype(np.float64), 5, 4)
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)


def test_coincident_and_misuse(ctx):
    # tests/IAgentForceBackend2DTests.cpp:144-165
    plan = halhip.Agents2D(ctx, 8, 8, 4)
    v = plan.accumulate_forces([4.0, 4.0, 4.0, 4.0], [0.5, 0.5], 2.0, 1.0, 0, 0)
    assert np.all(v == 0.0)
    with pytest.raises(halhip.InvalidArgument):
        plan.accumulate_forces([1.0, 1.0, 2.0], [0.5, 0.5], 2.0, 1.0, 0, 0)
    with pytest.raises(halhip.InvalidArgument):
        plan.accumulate_forces(np.ones(10), np.ones(5), 2.0, 1.0, 0, 0)   # over capacity
    empty = plan.accumulate_forces(np.zeros(0), np.zeros(0), 2.0, 1.0, 0, 0)
    assert empty.size == 0


@pytest.mark.parametrize("wrap", [(0, 0), (1, 1), (1, 0)])
def test_ticks_vs_oracle_small(ctx, oracle, wrap):
    rng = np.random.default_rng(3)
    X, Y, n = 64, 48, 2500
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, rad)
    plan.run_ticks(1, 1.0, 1.0, 0.1, *wrap)
    # One tick: bins of the initial positions are bit-exact; moved positions
    # agree to FP32 rounding.
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(pos.astype(np.float64), X, Y)
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)
    ref1 = oracle.agent_ticks(pos.astype(np.float64), rad.astype(np.float64), 1, 1.0, 1.0, 0.1, X, Y, *wrap)
    np.testing.assert_allclose(plan.positions(), ref1, rtol=0, atol=2e-5)
    plan.run_ticks(19, 1.0, 1.0, 0.1, *wrap)
    ref = oracle.agent_ticks(pos.astype(np.float64), rad.astype(np.float64), 20, 1.0, 1.0, 0.1, X, Y, *wrap)
    got = plan.positions().astype(np.float64)
    d = got - ref
    if wrap[0]:
        d[0::2] = (d[0::2] + X / 2) % X - X / 2
    if wrap[1]:
        d[1::2] = (d[1::2] + Y / 2) % Y - Y / 2
    assert np.abs(d).max() < 1e-3, np.abs(d).max()


def test_bins_each_tick_match_oracle_on_gpu_positions(ctx, oracle):
    # Per tick the GPU bins are bit-exact against the reference binning of the
    # GPU's own FP32 positions at that tick.
    rng = np.random.default_rng(5)
    X, Y, n = 128, 96, 10000
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, np.full(n, 0.5, np.float32))
    for _ in range(5):
        before = plan.positions()
        plan.run_ticks(1, 1.0, 1.0, 0.1, 0, 0)
        offsets, bucket = plan.bins()
        ro, rb = oracle.agent_bins(before.astype(np.float64), X, Y)
        np.testing.assert_array_equal(offsets, ro)
        np.testing.assert_array_equal(bucket, rb)


def test_workload_4m_first_tick(ctx, oracle):
    w = workloads.agents_uniform()
    n = w["rad"].size
    plan = halhip.Agents2D(ctx, w["X"], w["Y"], n)
    plan.upload(w["pos"], w["rad"])
    plan.run_ticks(0, w["R"], w["k"], 0.0, 0, 0)
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(w["pos"].astype(np.float64), w["X"], w["Y"])
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)
    v = plan.velocities().astype(np.float64)
    ref = oracle.agent_forces(w["pos"].astype(np.float64), w["rad"].astype(np.float64), w["R"], w["k"],
                              w["X"], w["Y"], 0, 0)
    err = rel_l2(v, ref)
    print(f"4M agents forces: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-5
