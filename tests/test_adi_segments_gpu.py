"""ADI rows longer than one workgroup (HALHIP_ADI_ROWS_SEGMENTED) against the
FP64 oracle's Thomas solves (Internal::Diffusion2DADI,
include/Hal/Pde/ADIequations.h:306-378, with TDMA / TDMAperiodic,
include/Hal/Pde/TdmaSolver.h:27-130, which has no cap on the line length).

Rows of yDim > HALHIP_ADI_MAX_Y are cut into 4096-column segments.  Forcing
the method on shorter rows uses 256-column segments, so small grids cross
many segment joins: every boundary kind, partial last segments, odd widths
(scalar rows) and rates whose line factors couple thousands of columns.

Tolerance: FP32 against FP64, relative L2 1e-5 (the bar of the whole-row
path); at r = 5000 the whole-row path itself sets the bar.
"""
import numpy as np
import pytest

import halhip
from helpers import rel_l2

pytestmark = pytest.mark.gpu

BOUNDARIES = ((0, 0, None), (1, 1, None), (1, 0, 0.3), (0, 1, 0.7), (0, 0, 0.0))


def ref_ticks(oracle, f, ticks, rate, X, Y, wx, wy, bc=None, decay=0.0, src_idx=(), src_val=()):
    return oracle.adi_ticks(np.asarray(f, np.float64), ticks, rate, decay, X, Y, wx, wy, bc=bc,
                            src_idx=src_idx, src_val=np.asarray(src_val, np.float64))


@pytest.mark.parametrize("X,Y", [(1, 257), (5, 300), (9, 1000), (33, 1025), (4, 2049), (40, 1024), (7, 8192)])
def test_forced_segments_all_boundaries(ctx, oracle, X, Y):
    rng = np.random.default_rng(X * 7919 + Y)
    f = rng.random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_methods(halhip.ADI_ROWS_SEGMENTED)
    for wx, wy, bc in BOUNDARIES:
        plan.upload(f)
        plan.run_steps(3, 5.0, 0.0, wx, wy, bc=bc)
        assert plan.methods()[0] == halhip.ADI_ROWS_SEGMENTED
        err = rel_l2(plan.download(), ref_ticks(oracle, f, 3, 5.0, X, Y, wx, wy, bc))
        assert err < 1e-5, (X, Y, wx, wy, bc, err)
    plan.close()


@pytest.mark.parametrize("rate", [0.05, 50.0, 5000.0])
def test_forced_segments_across_rates(ctx, oracle, rate):
    # At r = 5000 the multipliers are 0.98: a carry crosses several
    # 256-column segments before it fades.
    X, Y = 6, 3000
    rng = np.random.default_rng(int(rate * 100))
    f = rng.random(X * Y).astype(np.float32)
    seg = halhip.Adi2D(ctx, X, Y)
    seg.set_methods(halhip.ADI_ROWS_SEGMENTED)
    whole = halhip.Adi2D(ctx, X, Y)
    for wx, wy, bc in BOUNDARIES:
        ref = ref_ticks(oracle, f, 4, rate, X, Y, wx, wy, bc)
        errs = []
        for plan in (seg, whole):
            plan.upload(f)
            plan.run_steps(4, rate, 0.0, wx, wy, bc=bc)
            errs.append(rel_l2(plan.download(), ref))
        assert errs[0] < max(1e-5, 2.0 * errs[1]), (rate, wx, wy, bc, errs)
    seg.close()
    whole.close()


def test_forced_segments_eight_row_strips(ctx, oracle):
    X, Y = 37, 1200
    f = np.random.default_rng(3).random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_methods(halhip.ADI_ROWS_SEGMENTED)
    plan.set_strip_shape(4, 8, True)
    assert plan.strip_shape() == (4, 8, True)
    for wx, wy, bc in BOUNDARIES:
        plan.upload(f)
        plan.run_steps(3, 2.0, 0.0, wx, wy, bc=bc)
        err = rel_l2(plan.download(), ref_ticks(oracle, f, 3, 2.0, X, Y, wx, wy, bc))
        assert err < 1e-5, (wx, wy, bc, err)
    plan.close()


@pytest.mark.parametrize("X,Y", [(4, 12288), (3, 16384), (5, 16387), (1024, 16384)])
def test_long_rows_vs_oracle(ctx, oracle, X, Y):
    rng = np.random.default_rng(Y + X)
    f = rng.random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    bounds = BOUNDARIES if X * Y <= 1 << 20 else ((0, 0, 0.0), (1, 1, None))
    for wx, wy, bc in bounds:
        plan.upload(f)
        plan.run_steps(3, 5.0, 0.0, wx, wy, bc=bc)
        assert plan.methods()[0] == halhip.ADI_ROWS_SEGMENTED
        err = rel_l2(plan.download(), ref_ticks(oracle, f, 3, 5.0, X, Y, wx, wy, bc))
        print(f"ADI {X}x{Y} wrap {wx}{wy} bc {bc}: rel L2 vs FP64 = {err:.3e}")
        assert err < 1e-5, (X, Y, wx, wy, bc, err)
    plan.close()


def test_long_rows_decay_sources_and_graphs(ctx, oracle):
    # 45 model ticks (decay and point sources, Dirichlet x, periodic y):
    # two 20-tick graphs built before the run, then 5 direct ticks.
    X, Y = 12, 12288
    rng = np.random.default_rng(21)
    f = rng.random(X * Y).astype(np.float32)
    idx = [0, 4095, 4096, 4096, X * Y - 1, 5 * Y + 8191]
    val = [1.0, 0.5, 0.25, 0.25, 2.0, 0.125]
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    plan.set_sources(idx, val)
    plan.prepare_steps(45, 3.0, 4e-3, False, True, bc=0.2)
    plan.run_steps(45, 3.0, 4e-3, False, True, bc=0.2)
    err = rel_l2(plan.download(), ref_ticks(oracle, f, 45, 3.0, X, Y, 0, 1, 0.2, 4e-3, idx, val))
    print(f"ADI {X}x{Y}, 45 model ticks: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-5
    plan.close()


def test_long_rows_per_call_interface(ctx, oracle):
    # IAdiDiffusionBackend2D::DiffuseADI on a long row: deltas += ADI(u) - u.
    X, Y = 3, 9000
    f = np.random.default_rng(5).random(X * Y)
    plan = halhip.Adi2D(ctx, X, Y)
    d0 = np.full(X * Y, 0.5)
    got = plan.diffuse(f, d0.copy(), 0.8, True, False)
    want = d0 + ref_ticks(oracle, f, 1, 0.8, X, Y, 1, 0) - f
    np.testing.assert_allclose(got, want, rtol=0, atol=3e-6)
    plan.close()


def test_long_rows_method_limits(ctx):
    plan = halhip.Adi2D(ctx, 4, halhip.ADI_MAX_Y + 1)
    for rows, carries in ((halhip.ADI_ROWS_LDS, halhip.ADI_AUTO),
                          (halhip.ADI_ROWS_WORKGROUP_SCAN, halhip.ADI_AUTO),
                          (halhip.ADI_AUTO, halhip.ADI_CARRIES_WINDOW)):
        with pytest.raises(halhip.HalHipError) as e:
            plan.set_methods(rows, carries)
        assert e.value.code == halhip.ERR_UNSUPPORTED
    with pytest.raises(halhip.InvalidArgument):
        plan.set_strip_shape(8, 4, True)
    plan.set_methods(halhip.ADI_ROWS_SEGMENTED, halhip.ADI_CARRIES_SCAN)
    plan.close()
