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


@pytest.mark.parametrize("X,Y", [(5, 4), (7, 9), (3, 1)])
def test_crowded_cells_order(ctx, oracle, X, Y):
    # Many agents per cell exercise the per-cell ordering pass, which takes
    # four cells per thread: 63 and 3 cells leave a partial last group.
    rng = np.random.default_rng(7)
    n = 20000
    pos = (rng.random(2 * n) * np.tile([X + 2.0, Y + 2.0], n) - 1.0).astype(np.float32)   # some clamped
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, np.full(n, 0.1, np.float32))
    plan.run_ticks(0, 0.05, 1.0, 0.0, 0, 0)
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(pos.astype(np.float64), X, Y)
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


def test_hand_placed_pairs_closed_form(ctx):
    # tests/AgentForceKernel2DTests.cpp:49-123 and SphericalAgentTests.cpp:
    # linear overlap force k((ri + rj) - d) pushing each agent of a pair
    # apart along the pair's axis; none at d >= R; pairs across a periodic
    # edge by the minimum image.  Distances are exact in binary.
    k, R = 1.0, 1.0
    pos = [2.0, 3.0, 2.75, 3.0,      # pair along x, d = 0.75, overlap 0.25
           6.0, 6.0, 6.0, 6.5,       # pair along y, d = 0.5, radii 0.5 and 0.25: overlap 0.25
           10.0, 10.0, 11.5, 10.0,   # d = 1.5 >= R: no force
           0.25, 12.0, 15.5, 12.0]   # across the x edge of a 16-wide grid, d = 0.75
    rad = [0.5, 0.5, 0.5, 0.25, 0.5, 0.5, 0.5, 0.5]
    plan = halhip.Agents2D(ctx, 16, 16, 8)
    want = np.array([-0.25, 0.0, 0.25, 0.0,
                     0.0, -0.25, 0.0, 0.25,
                     0.0, 0.0, 0.0, 0.0,
                     0.25, 0.0, -0.25, 0.0])
    got = plan.accumulate_forces(pos, rad, R, k, 1, 0)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    # without wrap the last pair is 15.25 apart: no force
    want[12:] = 0.0
    got = plan.accumulate_forces(pos, rad, R, k, 0, 0)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)


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


def test_bins_on_a_grid_of_more_than_4096_scan_tiles(ctx, oracle):
    # 4305 scan tiles: the one-pass scan's look-back crosses many windows of
    # 64 predecessors.
    rng = np.random.default_rng(23)
    X, Y, n = 4200, 4100, 3000
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    pos[:200] = np.tile(pos[:2], 100)            # 100 agents in one cell
    rad = np.full(n, 0.5, np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, rad)
    plan.run_ticks(0, 1.0, 1.0, 0.0, 0, 0)
    offsets, bucket = plan.bins()
    ro, rb = oracle.agent_bins(pos.astype(np.float64), X, Y)
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)


def test_scan_wait_limit_fails_loudly_never_silently(ctx, oracle):
    # With a spin limit of 0 a scan whose predecessor tile has not yet
    # published gives up at once: the bins are then reported as failed, never
    # returned wrong.  Restoring the limit and binning again gives exact bins.
    rng = np.random.default_rng(29)
    X, Y, n = 2048, 2048, 200000
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    ro, rb = oracle.agent_bins(pos.astype(np.float64), X, Y)
    plan = halhip.Agents2D(ctx, X, Y, n)
    with pytest.raises(halhip.InvalidArgument):
        plan.set_scan_spin_limit(-1)
    plan.set_scan_spin_limit(0)
    outcomes = []
    for _ in range(6):
        plan.upload(pos, rad)
        plan.run_ticks(0, 1.0, 1.0, 0.0, 0, 0)
        try:
            offsets, bucket = plan.bins()
        except halhip.HalHipError as e:
            assert "look-back" in str(e)
            outcomes.append("failed")
            continue
        np.testing.assert_array_equal(offsets, ro)
        np.testing.assert_array_equal(bucket, rb)
        outcomes.append("exact")
    print("spin limit 0:", outcomes)
    plan.set_scan_spin_limit(1 << 20)
    plan.upload(pos, rad)
    plan.run_ticks(0, 1.0, 1.0, 0.0, 0, 0)
    offsets, bucket = plan.bins()
    np.testing.assert_array_equal(offsets, ro)
    np.testing.assert_array_equal(bucket, rb)
    plan.close()


def test_equal_radii_path_matches_general_path(ctx):
    # Equal radii take a force pass that loads no radius and a binning that
    # moves none.  The same agents plus one lone agent of another radius
    # (more than 4 cells from every other, beyond any drift in 12 ticks)
    # take the general path: every other agent's trajectory must be
    # bit-identical.  Re-uploads on one plan switch between the two and
    # between two equal radii over graph-replayed ticks.
    rng = np.random.default_rng(23)
    X, Y, n, ticks = 40, 36, 1300, 12
    pos = np.empty(2 * n, np.float32)
    pos[0::2] = rng.random(n) * 30.0
    pos[1::2] = rng.random(n) * Y
    pos_b = np.concatenate([pos, np.array([35.5, 18.0], np.float32)])
    plan = halhip.Agents2D(ctx, X, Y, n + 1)
    for r in (0.5, 0.3):
        rad = np.full(n, r, np.float32)
        rad_b = np.concatenate([rad, np.array([1.5 * r], np.float32)])
        plan.upload(pos, rad)
        plan.run_ticks(ticks, 1.0, 1.0, 0.1, 0, 1)
        equal = plan.positions()
        plan.upload(pos_b, rad_b)
        plan.run_ticks(ticks, 1.0, 1.0, 0.1, 0, 1)
        general = plan.positions()
        np.testing.assert_array_equal(equal, general[:2 * n])
        assert np.abs(general[2 * n:] - pos_b[2 * n:]).max() == 0.0   # the lone agent feels nothing
        plan.upload(pos, rad)
        plan.run_ticks(ticks, 1.0, 1.0, 0.1, 0, 1)
        np.testing.assert_array_equal(plan.positions(), equal)
    plan.close()


def test_graph_chunks_from_fresh_upload_and_forces_only(ctx, oracle):
    # The first graph chunk after an upload (or after a forces-only pass)
    # starts from the binned positions, later chunks from the moved ones.
    rng = np.random.default_rng(17)
    X, Y, n = 48, 40, 1500
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    ref = oracle.agent_ticks(pos.astype(np.float64), rad.astype(np.float64), 25, 1.0, 1.0, 0.1, X, Y, 0, 1)
    for _ in range(2):   # the second pass replays the cached graphs
        plan.upload(pos, rad)
        plan.run_ticks(25, 1.0, 1.0, 0.1, 0, 1)
        d = plan.positions().astype(np.float64) - ref
        d[1::2] = (d[1::2] + Y / 2) % Y - Y / 2
        assert np.abs(d).max() < 1e-3
    plan.upload(pos, rad)
    plan.run_ticks(0, 1.0, 1.0, 0.0, 0, 1)   # bins and forces only
    plan.run_ticks(20, 1.0, 1.0, 0.1, 0, 1)
    ref20 = oracle.agent_ticks(pos.astype(np.float64), rad.astype(np.float64), 20, 1.0, 1.0, 0.1, X, Y, 0, 1)
    d = plan.positions().astype(np.float64) - ref20
    d[1::2] = (d[1::2] + Y / 2) % Y - Y / 2
    assert np.abs(d).max() < 1e-3


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


@pytest.mark.parametrize("ticks", [20, 13, 1])
def test_velocities_are_the_last_ticks(ctx, oracle, ticks):
    # Resident ticks keep only the last tick's velocity deltas (a graph
    # chunk writes them on its last tick, an eager run on its last): after
    # run_ticks(n) they are the forces on the positions after n-1 ticks,
    # bit-identical to stepping n-1 ticks and then one more.
    rng = np.random.default_rng(41)
    X, Y, n = 48, 40, 2000
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, rad)
    plan.run_ticks(ticks, 1.0, 1.0, 0.1, 0, 1)
    v_run = plan.velocities()
    plan.upload(pos, rad)
    if ticks > 1:
        plan.run_ticks(ticks - 1, 1.0, 1.0, 0.1, 0, 1)
    before = plan.positions()
    plan.run_ticks(1, 1.0, 1.0, 0.1, 0, 1)
    np.testing.assert_array_equal(v_run, plan.velocities())
    ref = oracle.agent_forces(before.astype(np.float64), rad.astype(np.float64), 1.0, 1.0, X, Y, 0, 1)
    assert rel_l2(v_run, ref) < 1e-5
    plan.close()


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


# ---- carried velocity: friction and speed cap ------------------------------

def _wrapped_diff(got, ref, X, Y, wrap):
    d = got.astype(np.float64) - ref
    if wrap[0]:
        d[0::2] = (d[0::2] + X / 2) % X - X / 2
    if wrap[1]:
        d[1::2] = (d[1::2] + Y / 2) % Y - Y / 2
    return d


@pytest.mark.parametrize("wrap", [(0, 0), (1, 1)])
def test_motion_ticks_vs_oracle(ctx, oracle, wrap):
    # SphericalAgent2D-style ticks: carry + force -> cap -> move -> friction.
    rng = np.random.default_rng(31)
    X, Y, n = 64, 48, 2500
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    plan = halhip.Agents2D(ctx, X, Y, n)
    plan.upload(pos, rad)
    plan.set_motion(0.7, 1.5)
    plan.run_ticks(1, 1.0, 1.0, 0.1, *wrap)
    p1, c1 = oracle.agent_ticks_motion(pos, np.zeros(2 * n), rad, 1, 1.0, 1.0, 0.1, 0.7, 1.5, X, Y, *wrap)
    np.testing.assert_allclose(plan.positions(), p1, rtol=0, atol=2e-5)
    np.testing.assert_allclose(plan.motion(), c1, rtol=0, atol=2e-5)
    plan.run_ticks(19, 1.0, 1.0, 0.1, *wrap)    # a 10-tick graph chunk and single ticks
    p, c = oracle.agent_ticks_motion(pos, np.zeros(2 * n), rad, 20, 1.0, 1.0, 0.1, 0.7, 1.5, X, Y, *wrap)
    assert np.abs(_wrapped_diff(plan.positions(), p, X, Y, wrap)).max() < 1e-3
    assert np.abs(plan.motion() - c).max() < 1e-3
    # the cap holds: |carry| <= friction * maxVel
    cm = plan.motion().reshape(-1, 2)
    assert np.sqrt((cm.astype(np.float64) ** 2).sum(1)).max() <= 0.7 * 1.5 * (1 + 1e-6)


def test_motion_closed_form_and_checkpoint(ctx, oracle):
    # Lone agents: ApplyFriction(0.5) / CapVelocity(2) on (3, 4)
    # (tests/SphericalAgentTests.cpp:88-102), one tick with dt = 1.
    plan = halhip.Agents2D(ctx, 40, 40, 2)
    plan.upload(np.array([10.0, 10.0, 30.0, 30.0], np.float32), np.array([0.5, 0.5], np.float32))
    plan.set_motion(0.5, 2.0)
    plan.upload_motion(np.array([3.0, 4.0, -0.3, 0.4], np.float32))
    plan.run_ticks(1, 1.0, 1.0, 1.0, 0, 0)
    np.testing.assert_allclose(plan.positions(), [11.2, 11.6, 29.7, 30.4], rtol=0, atol=1e-5)
    np.testing.assert_allclose(plan.motion(), [0.6, 0.8, -0.15, 0.2], rtol=0, atol=1e-6)

    # Checkpoint (positions, radii, carried velocities) and resume on a new
    # plan: bins are stable by agent index, so the resumed run is bit-exact.
    rng = np.random.default_rng(32)
    X, Y, n = 40, 30, 900
    pos = (rng.random(2 * n) * np.tile([X, Y], n)).astype(np.float32)
    rad = np.full(n, 0.5, np.float32)
    a = halhip.Agents2D(ctx, X, Y, n)
    a.upload(pos, rad)
    a.set_motion(0.8, 0.0)
    a.run_ticks(7, 1.0, 1.0, 0.2, 0, 1)
    ckpt_pos, ckpt_vel = a.positions(), a.motion()
    a.run_ticks(13, 1.0, 1.0, 0.2, 0, 1)
    b = halhip.Agents2D(ctx, X, Y, n)
    b.upload(ckpt_pos, rad)
    b.set_motion(0.8, 0.0)
    b.upload_motion(ckpt_vel)
    b.run_ticks(13, 1.0, 1.0, 0.2, 0, 1)
    np.testing.assert_array_equal(b.positions(), a.positions())
    np.testing.assert_array_equal(b.motion(), a.motion())


def test_motion_misuse_and_reset(ctx):
    plan = halhip.Agents2D(ctx, 8, 8, 4)
    plan.upload(np.array([1.0, 1.0, 5.0, 5.0], np.float32), np.array([0.5, 0.5], np.float32))
    with pytest.raises(halhip.InvalidArgument):
        plan.upload_motion(np.zeros(4, np.float32))      # no friction or cap set
    with pytest.raises(halhip.InvalidArgument):
        plan.set_motion(1.5, 0.0)
    with pytest.raises(halhip.InvalidArgument):
        plan.set_motion(0.5, -1.0)
    plan.set_motion(0.9, 0.0)
    plan.upload_motion(np.array([0.5, 0.0, 0.0, 0.5], np.float32))
    plan.upload(np.array([1.0, 1.0, 5.0, 5.0], np.float32), np.array([0.5, 0.5], np.float32))
    assert np.all(plan.motion() == 0.0)                  # an upload starts agents at rest
    plan.set_motion(0.0, 0.0)                            # off: plain ticks again
    plan.run_ticks(2, 1.0, 1.0, 0.1, 0, 0)
    np.testing.assert_array_equal(plan.positions(), np.array([1.0, 1.0, 5.0, 5.0], np.float32))
