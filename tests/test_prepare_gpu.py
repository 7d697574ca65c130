"""Graph preparation (halhip_*_prepare_steps / _prepare_ticks): a prepare call
captures the graphs a run replays and runs no tick, so the resident state is
unchanged by it, and a run after it gives the same bits as the same run
without it (the same kernels with the same arguments, replayed or captured on
first use)."""
import numpy as np
import pytest

import halhip
import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("X,Y,decay", [(96, 512, 0.0), (64, 256, 1e-3), (4096, 4096, 0.0)])
def test_adi_prepare_keeps_state_and_bits(ctx, X, Y, decay):
    f = np.random.default_rng(X + Y).random(X * Y).astype(np.float32)
    runs = []
    for prepare in (False, True):
        plan = halhip.Adi2D(ctx, X, Y)
        plan.set_sources([7, X * Y // 2], [0.5, 0.25])
        plan.upload(f)
        plan.run_steps(5, 5.0, decay, False, False, bc=0.0)  # leaves the odd reduction half
        if prepare:
            before = plan.download()
            plan.prepare_steps(40, 5.0, decay, False, False, bc=0.0)
            assert np.array_equal(plan.download(), before)
        plan.run_steps(40, 5.0, decay, False, False, bc=0.0)
        runs.append(plan.download())
        plan.close()
    print(f"ADI {X}x{Y} decay {decay}: prepared run bit-identical", flush=True)
    assert np.array_equal(runs[0], runs[1])


def test_explicit_prepare_keeps_state_and_bits(ctx):
    w = workloads.explicit_256()
    runs = []
    for prepare in (False, True):
        p = halhip.Diffusion2D(ctx, w["X"], w["Y"])
        p.set_sources(w["src_idx"], w["src_val"])
        p.upload(w["field"])
        p.run_steps(3, w["rate"], w["decay"], True, True)
        if prepare:
            before = p.download()
            p.prepare_steps(200, w["rate"], w["decay"], True, True)
            assert np.array_equal(p.download(), before)
        p.run_steps(200, w["rate"], w["decay"], True, True)
        runs.append(p.download())
        p.close()
    assert np.array_equal(runs[0], runs[1])


def test_agents_prepare_keeps_state_and_bits(ctx):
    w = workloads.agents_uniform(n_agents=20_000, dim=160)
    runs = []
    for prepare in (False, True):
        p = halhip.Agents2D(ctx, w["X"], w["Y"], w["rad"].size)
        p.upload(w["pos"], w["rad"])
        if prepare:
            p.prepare_ticks(30, w["R"], w["k"], w["dt"], False, False)
            assert np.array_equal(p.positions(), w["pos"].astype(np.float32))
        p.run_ticks(30, w["R"], w["k"], w["dt"], False, False)
        runs.append((p.positions(), p.velocities()))
        p.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_batch_prepare_keeps_state_and_bits(ctx):
    spec = workloads.wells_batch(n_wells=12, seed=5, hi=192)
    runs = []
    for prepare in (False, True):
        b = halhip.Batch(ctx, spec["xdims"], spec["ydims"])
        b.set_wells(spec["rate"], spec["decay"])
        b.set_sources(spec["src_well"], spec["src_cell"], spec["src_val"])
        b.upload(np.zeros(b.cells, np.float32))
        b.run_ticks(3)
        if prepare:
            before = b.download()
            b.prepare_ticks(30)
            assert np.array_equal(b.download(), before)
        b.run_ticks(30)
        runs.append(b.download())
        b.close()
    assert np.array_equal(runs[0], runs[1])
