"""ADI on the GPU vs the FP64 oracle, the reference's golden vectors and its
closed forms (IAdiDiffusionBackend2D / AdiKernel2D::RunSteps /
PDEGrid2D::DiffusionADI with Dirichlet values).

Tolerances: FP32 arithmetic against FP64; the relative L2 bar is 1e-5.
"""
import numpy as np
import pytest

import halhip
import workloads
from helpers import eigen_field, golden_cases, lambda_cn, rel_l2

pytestmark = pytest.mark.gpu


def adi_ref_ticks(oracle, f, ticks, rate, X, Y, wx, wy, bc=None, decay=0.0, src_idx=(), src_val=()):
    return oracle.adi_ticks(np.asarray(f, np.float64), ticks, rate, decay, X, Y, wx, wy, bc=bc,
                            src_idx=src_idx, src_val=np.asarray(src_val, np.float64))


def test_interface_matches_golden_cases(ctx, golden):
    for c in golden_cases(golden):
        if not c["adi"]:
            continue
        plan = halhip.Adi2D(ctx, c["X"], c["Y"])
        got = plan.diffuse(c["field"], c["deltas_in"].copy(), c["rate"], c["wrapX"], c["wrapY"], c["bc"])
        scale = max(1.0, np.abs(c["field"]).max())
        np.testing.assert_allclose(got, c["deltas_out"], rtol=0, atol=3e-6 * scale,
                                   err_msg=f"case {c['k']} X={c['X']} Y={c['Y']} wrap={c['wrapX']}{c['wrapY']}")
        plan.close()


def test_eigenmode_through_interface(ctx):
    # tests/IDiffusionBackend2DTests.cpp:76-144 (1e-4 there), called twice on
    # one plan like the reference test.
    nx, ny, r = 12, 18, 0.3
    f = eigen_field(nx, ny)
    plan = halhip.Adi2D(ctx, nx, ny)
    lam = lambda_cn(r, nx) * lambda_cn(r, ny)
    for _ in range(2):
        got = f + plan.diffuse(f, np.zeros_like(f), r, True, True)
        np.testing.assert_allclose(got, lam * f, rtol=0, atol=2e-6)


def test_per_call_leaves_resident_field(ctx, oracle):
    # A per-call DiffuseADI between resident runs changes neither the resident
    # field nor the ticks that follow it: 3 + 4 resident ticks equal 7 oracle
    # ticks, and the call's own deltas match the oracle's one tick.
    X, Y, r = 40, 512, 2.0
    rng = np.random.default_rng(11)
    f = rng.random(X * Y).astype(np.float32)
    g = rng.random(X * Y)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    plan.run_steps(3, r, 0.0, False, False, bc=0.0)
    d = plan.diffuse(g, np.zeros_like(g), r, False, False, 0.0)
    plan.run_steps(4, r, 0.0, False, False, bc=0.0)
    got = plan.download()
    ref = adi_ref_ticks(oracle, f, 7, r, X, Y, 0, 0, bc=0.0)
    assert rel_l2(got, ref) < 1e-5
    want = adi_ref_ticks(oracle, g, 1, r, X, Y, 0, 0, bc=0.0) - g
    np.testing.assert_allclose(d, want, rtol=0, atol=3e-6)
    plan.close()


def test_resident_eigenmode_powers(ctx):
    # tests/AdiKernel2DTests.cpp:18-68: the periodic eigenmode after N = 50
    # resident ticks at r = 0.8 is (lam_x lam_y)^N f, then ^2N (1e-3 there).
    # The mode decays 5000-fold over 2N ticks while the FP32 rounding of the
    # first ticks does not, so the error relative to the decayed mode grows
    # to ~3e-5; relative to the initial field it stays ~1e-8.
    nx, ny, r, n = 24, 32, 0.8, 50
    f = eigen_field(nx, ny)
    lam = lambda_cn(r, nx) * lambda_cn(r, ny)
    plan = halhip.Adi2D(ctx, nx, ny)
    plan.upload(f)
    for k, bar in ((1, 1e-5), (2, 1e-4)):
        plan.run_steps(n, r, 0.0, True, True)
        got = plan.download()
        want = lam ** (k * n) * f
        assert rel_l2(got, want) < bar
        assert np.linalg.norm(got - want) < 1e-7 * np.linalg.norm(f)


def test_zero_flux_mass_and_cell_for_cell(ctx, oracle):
    # tests/AdiKernel2DTests.cpp:70-141: 20x14, 5 steps, r = 0.6
    nx, ny = 20, 14
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    f = (1.0 + 0.3 * x - 0.2 * y + 0.5 * np.sin(0.7 * x * y)).ravel().astype(np.float32)
    plan = halhip.Adi2D(ctx, nx, ny)
    plan.upload(f)
    plan.run_steps(5, 0.6, 0.0, False, False)
    got = plan.download()
    ref = adi_ref_ticks(oracle, f, 5, 0.6, nx, ny, 0, 0)
    assert abs(got.astype(np.float64).sum() - f.astype(np.float64).sum()) < 1e-4 * abs(f.sum())
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)


def test_dirichlet_fixed_point_relaxation_and_ramp(ctx):
    # tests/PDEADITests.cpp:219-264, 297-354
    plan = halhip.Adi2D(ctx, 7, 5)
    plan.upload(np.full(35, 4.0, np.float32))
    plan.run_steps(3, 0.9, 0.0, False, False, bc=4.0)
    np.testing.assert_allclose(plan.download(), 4.0, rtol=1e-6)

    plan = halhip.Adi2D(ctx, 8, 6)
    plan.upload(np.zeros(48, np.float32))
    plan.run_steps(400, 1.5, 0.0, True, False, bc=2.0)
    np.testing.assert_allclose(plan.download(), 2.0, rtol=1e-5)

    X, Y = 6, 4
    ramp = lambda x: 1.0 if x < 0 else (3.0 if x >= X else 1.0 + 2.0 * (x + 0.5) / X)
    f = np.array([ramp(x) for x in range(X) for _ in range(Y)])
    edges = (np.full(Y, 1.0), np.full(Y, 3.0), np.array([ramp(x) for x in range(X)]),
             np.array([ramp(x) for x in range(X)]))
    plan = halhip.Adi2D(ctx, X, Y)
    d = plan.diffuse(f, np.zeros_like(f), 0.9, False, False, edges)
    np.testing.assert_allclose(d, 0.0, atol=2e-6)


# Shapes that exercise every strip kernel variant (columns per thread 1..8,
# strip heights 4..16, partial last strips and wavefronts, odd widths).
SHAPES = [(1, 1), (3, 1), (1, 7), (5, 64), (9, 65), (33, 200), (70, 1000), (17, 1025), (13, 2048), (21, 2049),
          (11, 4096), (6, 5000), (5, 8192), (300, 300), (1030, 40)]


@pytest.mark.parametrize("X,Y", SHAPES)
def test_shapes_all_boundaries_vs_oracle(ctx, oracle, X, Y):
    rng = np.random.default_rng(X * 10007 + Y)
    f = rng.random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    for wx, wy, bc in ((0, 0, None), (1, 1, None), (1, 0, 0.3), (0, 1, 0.7), (0, 0, 0.0)):
        plan.upload(f)
        plan.run_steps(3, 5.0, 0.0, wx, wy, bc=bc)
        got = plan.download()
        ref = adi_ref_ticks(oracle, f, 3, 5.0, X, Y, wx, wy, bc)
        err = rel_l2(got, ref)
        assert err < 1e-5, (X, Y, wx, wy, bc, err)


def test_decay_and_sources_vs_oracle(ctx, oracle):
    X, Y = 50, 130
    rng = np.random.default_rng(9)
    f = rng.random(X * Y).astype(np.float32)
    idx = [0, 77, 77, X * Y - 1, 2500]
    val = [1.0, 0.5, 0.25, 2.0, 0.125]
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    plan.set_sources(idx, val)
    plan.run_steps(25, 2.0, 5e-3, False, True, bc=0.1)
    got = plan.download()
    ref = adi_ref_ticks(oracle, f, 25, 2.0, X, Y, 0, 1, 0.1, 5e-3, idx, val)
    assert rel_l2(got, ref) < 1e-5


def test_rate_change_between_runs(ctx, oracle):
    X, Y = 40, 48
    f = np.random.default_rng(4).random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    plan.run_steps(4, 5.0, 0.0, False, False)
    plan.run_steps(4, 0.5, 0.0, False, False, bc=1.0)
    plan.run_steps(30, 5.0, 0.0, False, False)
    got = plan.download()
    ref = adi_ref_ticks(oracle, f, 4, 5.0, X, Y, 0, 0)
    ref = adi_ref_ticks(oracle, ref, 4, 0.5, X, Y, 0, 0, 1.0)
    ref = adi_ref_ticks(oracle, ref, 30, 5.0, X, Y, 0, 0)
    assert rel_l2(got, ref) < 1e-5


def test_errors(ctx):
    plan = halhip.Adi2D(ctx, 3, 4)
    with pytest.raises(halhip.InvalidArgument):
        plan.diffuse(np.zeros(12), np.zeros(11), 0.1, False, False)
    ctx.fail_next_allocs(1)
    with pytest.raises(halhip.HalHipError) as e:
        halhip.Adi2D(ctx, 64, 64)
    assert e.value.code == halhip.ERR_OUT_OF_MEMORY


# ---- the 4096^2 workload ------------------------------------------------------

def test_blobs_4096_first_ticks_vs_oracle(ctx, oracle):
    w = workloads.adi_blobs()
    plan = halhip.Adi2D(ctx, w["X"], w["Y"])
    plan.upload(w["field"])
    plan.run_steps(2, w["rate"], 0.0, False, False, bc=w["bc"])
    got = plan.download()
    ref = adi_ref_ticks(oracle, w["field"], 2, w["rate"], w["X"], w["Y"], 0, 0, w["bc"])
    err = rel_l2(got, ref)
    print(f"ADI 4096^2 blobs, 2 ticks: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-5


def test_4096_periodic_eigenmode_200_ticks(ctx):
    # Closed form at full size: each tick multiplies the mode by the
    # Crank-Nicolson factor product.
    n, r, k = 4096, 5.0, 3
    x = np.arange(n)
    col = np.sin(2 * np.pi * k * x / n)
    f = np.outer(col, col).ravel().astype(np.float32)
    plan = halhip.Adi2D(ctx, n, n)
    plan.upload(f)
    plan.run_steps(200, r, 0.0, True, True)
    got = plan.download()
    lam = (lambda_cn(r, n, k) ** 2) ** 200
    err = rel_l2(got, lam * f.astype(np.float64))
    print(f"4096^2 periodic eigenmode x 200 ticks: rel L2 = {err:.3e}")
    assert err < 1e-5


def test_4096_zero_flux_mass_200_ticks(ctx):
    w = workloads.adi_blobs()
    plan = halhip.Adi2D(ctx, w["X"], w["Y"])
    plan.upload(w["field"])
    plan.run_steps(200, w["rate"], 0.0, False, False)
    got = plan.download().astype(np.float64)
    m0 = w["field"].astype(np.float64).sum()
    assert abs(got.sum() - m0) < 1e-5 * m0


def test_4096_linearity(ctx):
    n = 4096
    w = workloads.adi_blobs()
    g = np.random.default_rng(1).random(n * n).astype(np.float32)
    plan = halhip.Adi2D(ctx, n, n)
    outs = []
    for f in (w["field"], g, 2.0 * w["field"] - 3.0 * g):
        plan.upload(f)
        plan.run_steps(20, 5.0, 0.0, False, False, bc=0.0)
        outs.append(plan.download().astype(np.float64))
    assert rel_l2(outs[2], 2.0 * outs[0] - 3.0 * outs[1]) < 1e-5


@pytest.mark.parametrize("n,shape", [(512, (1, 4, True)), (1024, (2, 4, True)), (2048, (4, 8, True)),
                                     (4096, (4, 8, True)), (1000, (1, 4, True))])
def test_automatic_shape(ctx, n, shape):
    # Up to yDim 2048 the automatic shape prefers 512-thread workgroups
    # where the LDS row solve covers them; 1000 columns (not a multiple of
    # 64) keep the register solve at one column per thread.
    plan = halhip.Adi2D(ctx, n, n)
    assert plan.strip_shape() == shape
    plan.upload(np.zeros(n * n, np.float32))
    plan.run_steps(1, 5.0, 0.0, False, False, bc=0.0)
    want = halhip.ADI_ROWS_LDS if n % 64 == 0 else halhip.ADI_ROWS_WORKGROUP_SCAN
    assert plan.methods()[0] == want


@pytest.mark.parametrize("Y", [4096, 1000])
def test_every_strip_shape_matches_oracle(ctx, oracle, Y):
    X = 150
    f = np.random.default_rng(Y).random(X * Y).astype(np.float32)
    ref = {}
    plan = halhip.Adi2D(ctx, X, Y)
    for cpt in (1, 2, 4, 8):
        if (Y + cpt - 1) // cpt > 1024:
            continue
        for s in ((4, 8) if cpt == 8 else (4, 8, 16)):
            for vec in ((True, False) if Y % cpt == 0 else (False,)):
                plan.set_strip_shape(cpt, s, vec)
                assert plan.strip_shape() == (cpt, s, vec)
                for wx, wy, bc in ((0, 0, 0.0), (1, 1, None)):
                    plan.upload(f)
                    plan.run_steps(3, 5.0, 0.0, wx, wy, bc=bc)
                    key = (wx, wy, bc)
                    if key not in ref:
                        ref[key] = adi_ref_ticks(oracle, f, 3, 5.0, X, Y, wx, wy, bc)
                    err = rel_l2(plan.download(), ref[key])
                    assert err < 1e-5, (cpt, s, vec, key, err)


# ---- plain-tick methods: workgroup-scan vs LDS rows, scan vs windowed carries ----

METHOD_CASES = [
    # X, Y, (cpt, rows), rate, wrapX, bc
    (300, 1024, (1, 8), 5.0, 0, 0.2),
    (256, 2048, (2, 8), 5.0, 0, None),
    (200, 4096, (4, 8), 0.5, 0, 1.0),
    (200, 4096, (4, 8), 50.0, 0, 0.0),
    (96, 512, (1, 8), 5.0, 0, None),
    (128, 4096, (4, 8), 5.0, 1, None),
    (61, 4096, (4, 4), 2.0, 0, 0.5),
    (70, 2048, (2, 4), 5.0, 0, None),
    (40, 8192, (8, 4), 5.0, 0, 0.0),
    (100, 4096, (8, 4), 5.0, 0, 0.0),
    (90, 2048, (4, 8), 1.0, 0, None),
    (2, 1024, (1, 8), 5.0, 0, 0.0),
]


@pytest.mark.parametrize("X,Y,shape,rate,wx,bc", METHOD_CASES)
def test_methods_match_oracle(ctx, oracle, X, Y, shape, rate, wx, bc):
    f = np.random.default_rng(X + Y).random(X * Y).astype(np.float32)
    ref = adi_ref_ticks(oracle, f, 6, rate, X, Y, wx, 0, bc)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_strip_shape(shape[0], shape[1], True)
    used = set()
    for rows, carries in ((halhip.ADI_ROWS_WORKGROUP_SCAN, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW)):
        plan.set_methods(rows, carries)
        plan.upload(f)
        plan.run_steps(6, rate, 0.0, wx, 0, bc=bc)
        err = rel_l2(plan.download(), ref)
        m = plan.methods()
        used.add(m[:2])
        assert err < 1e-5, (rows, carries, m, err)
        # run 2 more ticks after a pause: the resident reductions carry over
        plan.run_steps(2, rate, 0.0, wx, 0, bc=bc)
    # the LDS row solve applies to every case here; windowed carries to the
    # non-periodic ones whose factors decay fast enough
    assert (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_SCAN) in used
    if not wx and rate <= 5.0:
        assert (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW) in used


PERIODIC_Y_CASES = [
    # X, Y, (cpt, rows), rate, wrapX, bc (Dirichlet applies to x only)
    (300, 1024, (1, 8), 5.0, 0, 0.2),
    (200, 4096, (4, 8), 5.0, 0, None),
    (128, 4096, (4, 8), 5.0, 1, None),
    (61, 4096, (4, 4), 50.0, 0, 0.5),
    (100, 4096, (8, 4), 0.5, 1, None),
    (96, 512, (1, 4), 5.0, 0, None),
]


@pytest.mark.parametrize("X,Y,shape,rate,wx,bc", PERIODIC_Y_CASES)
def test_periodic_y_lds_rows_match_oracle(ctx, oracle, X, Y, shape, rate, wx, bc):
    # A periodic y axis solves each row as the cyclic system of TDMAperiodic
    # in LDS (a rank-2 correction with off-diagonal corners), against the
    # register solve and the FP64 oracle.
    f = np.random.default_rng(X * 7 + Y).random(X * Y).astype(np.float32)
    ref = adi_ref_ticks(oracle, f, 6, rate, X, Y, wx, 1, bc)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_strip_shape(shape[0], shape[1], True)
    used = set()
    for rows, carries in ((halhip.ADI_ROWS_WORKGROUP_SCAN, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW)):
        plan.set_methods(rows, carries)
        plan.upload(f)
        plan.run_steps(6, rate, 0.0, wx, 1, bc=bc)
        m = plan.methods()
        used.add(m[0])
        err = rel_l2(plan.download(), ref)
        assert err < 1e-5, (rows, carries, m, err)
    assert halhip.ADI_ROWS_LDS in used


@pytest.mark.parametrize("rate", [0.05, 0.5, 5.0])
@pytest.mark.parametrize("wy", [0, 1])
def test_periodic_x_windowed_carries_match_oracle(ctx, oracle, rate, wy):
    # A periodic x axis with windowed carries: every strip forms the
    # Sherman-Morrison column values from the first and last strips'
    # reductions (no scan kernel).  Those windows start at the grid's ends,
    # where the strip multipliers are not zero, so they can be longer than
    # the carries' window; 1000 rows in 250 strips truncate both.
    X, Y = 1000, 1024
    f = np.random.default_rng(int(rate * 100) + wy).random(X * Y).astype(np.float32)
    ref = adi_ref_ticks(oracle, f, 4, rate, X, Y, 1, wy, None)
    plan = halhip.Adi2D(ctx, X, Y)
    got = {}
    for carries in (halhip.ADI_CARRIES_SCAN, halhip.ADI_CARRIES_WINDOW):
        plan.set_methods(halhip.ADI_ROWS_LDS, carries)
        plan.upload(f)
        plan.run_steps(4, rate, 0.0, 1, wy)
        got[plan.methods()[1]] = plan.download()
    assert set(got) == {halhip.ADI_CARRIES_SCAN, halhip.ADI_CARRIES_WINDOW}
    for carries, d in got.items():
        assert rel_l2(d, ref) < 1e-5, (carries, rel_l2(d, ref))
    assert rel_l2(got[halhip.ADI_CARRIES_WINDOW], got[halhip.ADI_CARRIES_SCAN]) < 2e-6


DECAY_CASES = [
    # X, Y, (cpt, rows), rate, wrapX, bc
    (300, 1024, (1, 8), 5.0, 0, 0.2),
    (200, 4096, (4, 8), 5.0, 0, 0.0),
    (61, 4096, (4, 4), 2.0, 0, 0.5),
    (100, 4096, (8, 4), 5.0, 0, None),
    (128, 2048, (4, 8), 5.0, 1, None),
]


@pytest.mark.parametrize("X,Y,shape,rate,wx,bc", DECAY_CASES)
def test_decay_and_sources_every_method_match_oracle(ctx, oracle, X, Y, shape, rate, wx, bc):
    # DiffusionADI + MulAll(-k) + Add + Update ticks: on the LDS row solve the
    # decay and the strip's sources are applied inside the tick kernel.
    # Sources: a repeated cell, the first and last cells, both sides of a
    # strip boundary; runs of odd length and graph-sized chunks.
    S = shape[1]
    f = np.random.default_rng(X * 3 + Y).random(X * Y).astype(np.float32)
    idx = [0, 5 * Y + 7, 5 * Y + 7, X * Y - 1, (S - 1) * Y + 100, S * Y + 100, (X // 2) * Y + Y - 1]
    val = [1.0, 0.5, 0.25, 2.0, 0.75, -0.5, 0.125]
    ref = adi_ref_ticks(oracle, f, 25, rate, X, Y, wx, 0, bc, 4e-3, idx, val)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_strip_shape(shape[0], shape[1], True)
    plan.set_sources(idx, val)
    used = set()
    for rows, carries in ((halhip.ADI_ROWS_WORKGROUP_SCAN, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_SCAN),
                          (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW)):
        plan.set_methods(rows, carries)
        plan.upload(f)
        for n in (1, 21, 3):
            plan.run_steps(n, rate, 4e-3, wx, 0, bc=bc)
        m = plan.methods()
        used.add(m[:2])
        err = rel_l2(plan.download(), ref)
        assert err < 1e-5, (rows, carries, m, err)
    assert (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_SCAN) in used


@pytest.mark.parametrize("X,wx", [(9000, 1), (9000, 0), (3000, 1)])
def test_tall_grids_scan_strip_counts(ctx, oracle, X, wx):
    # More than 1024 strips: the strip scan reloads its values instead of
    # keeping 16 per thread (9000 rows); 3000 rows keep them in registers.
    Y = 64
    f = np.random.default_rng(X + wx).random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_methods(halhip.ADI_ROWS_WORKGROUP_SCAN, halhip.ADI_CARRIES_SCAN)
    plan.upload(f)
    plan.run_steps(3, 5.0, 0.0, wx, 1, bc=None)
    ref = adi_ref_ticks(oracle, f, 3, 5.0, X, Y, wx, 1)
    assert rel_l2(plan.download(), ref) < 1e-5


@pytest.mark.parametrize("wx", [0, 1])
def test_periodic_y_model_ticks(ctx, oracle, wx):
    # Decay and point sources in the LDS tick on a periodic y axis.
    X, Y = 120, 2048
    f = np.random.default_rng(31 + wx).random(X * Y).astype(np.float32)
    idx = [0, Y - 1, 7 * Y + 100, 7 * Y + 100, X * Y - 1]
    val = [1.0, 0.5, 0.25, 0.25, 2.0]
    plan = halhip.Adi2D(ctx, X, Y)
    plan.set_sources(idx, val)
    plan.upload(f)
    for n in (1, 20, 4):
        plan.run_steps(n, 5.0, 3e-3, wx, 1, bc=None)
    assert plan.methods()[0] == halhip.ADI_ROWS_LDS
    ref = adi_ref_ticks(oracle, f, 25, 5.0, X, Y, wx, 1, None, 3e-3, idx, val)
    assert rel_l2(plan.download(), ref) < 1e-5


def test_plain_and_decay_runs_chain(ctx, oracle):
    # The decay ticks leave the reduction of their result for a following
    # plain run, and the other way round.
    X, Y = 96, 4096
    f = np.random.default_rng(8).random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    plan.run_steps(3, 5.0, 0.0, False, False, bc=0.0)
    plan.set_sources([3 * Y + 9], [1.5])
    plan.run_steps(22, 5.0, 1e-2, False, False, bc=0.0)
    plan.set_sources([], [])
    plan.run_steps(5, 5.0, 0.0, False, False, bc=0.0)
    ref = adi_ref_ticks(oracle, f, 3, 5.0, X, Y, 0, 0, 0.0)
    ref = adi_ref_ticks(oracle, ref, 22, 5.0, X, Y, 0, 0, 0.0, 1e-2, [3 * Y + 9], [1.5])
    ref = adi_ref_ticks(oracle, ref, 5, 5.0, X, Y, 0, 0, 0.0)
    assert plan.methods()[0] == halhip.ADI_ROWS_LDS
    assert rel_l2(plan.download(), ref) < 1e-5


def test_methods_odd_tick_counts_keep_reductions(ctx, oracle):
    # Windowed ticks alternate two reduction buffers; runs of odd length and
    # graph-sized chunks must chain correctly.
    X, Y = 160, 4096
    f = np.random.default_rng(5).random(X * Y).astype(np.float32)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(f)
    for n in (1, 3, 20, 21, 2):
        plan.run_steps(n, 5.0, 0.0, False, False, bc=0.0)
    assert plan.methods() == (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW, plan.methods()[2])
    ref = adi_ref_ticks(oracle, f, 47, 5.0, X, Y, 0, 0, 0.0)
    assert rel_l2(plan.download(), ref) < 1e-5


def test_blobs_4096_uses_lds_rows_and_windowed_carries(ctx):
    w = workloads.adi_blobs()
    plan = halhip.Adi2D(ctx, w["X"], w["Y"])
    plan.upload(w["field"])
    plan.run_steps(1, w["rate"], 0.0, False, False, bc=w["bc"])
    rows, carries, k = plan.methods()
    assert (rows, carries) == (halhip.ADI_ROWS_LDS, halhip.ADI_CARRIES_WINDOW) and 1 <= k <= 8
