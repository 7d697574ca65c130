"""Explicit diffusion on the GPU vs the FP64 oracle and the reference's
closed forms (IDiffusionBackend2D / DiffusionKernel2D::RunSteps)."""
import numpy as np
import pytest

import halhip
import workloads
from helpers import eigen_field, golden_cases, rel_l2

pytestmark = pytest.mark.gpu


def test_interface_matches_golden_cases(ctx, golden):
    for c in golden_cases(golden):
        if c["adi"]:
            continue
        plan = halhip.Diffusion2D(ctx, c["X"], c["Y"])
        # The interface computes in FP32 like the reference's GPU adapter:
        # compare against the reference on the FP32-rounded field.
        got = plan.diffuse(c["field"], c["deltas_in"].copy(), c["rate"], c["wrapX"], c["wrapY"], c["bc"])
        scale = max(1.0, np.abs(c["field"]).max())
        np.testing.assert_allclose(got, c["deltas_out"], rtol=0, atol=2e-6 * scale, err_msg=f"case {c['k']}")
        plan.close()


def test_eigenmode_through_interface(ctx):
    # tests/IDiffusionBackend2DTests.cpp:23-72 (tolerance 1e-4 relative there)
    nx, ny, r = 16, 20, 0.05
    f = eigen_field(nx, ny)
    plan = halhip.Diffusion2D(ctx, nx, ny)
    got = f + plan.diffuse(f, np.zeros_like(f), r, True, True)
    lam = 1 - 4 * r * (np.sin(np.pi / nx) ** 2 + np.sin(np.pi / ny) ** 2)
    np.testing.assert_allclose(got, lam * f, rtol=0, atol=1e-6)


def test_resident_eigenmode_powers(ctx):
    # tests/DiffusionKernel2DTests.cpp:16-107: after N and 2N resident steps
    # the periodic eigenmode is lam^N f and lam^2N f (1e-3 there).
    nx, ny, r, n = 16, 20, 0.05, 50
    f = eigen_field(nx, ny)
    lam = 1 - 4 * r * (np.sin(np.pi / nx) ** 2 + np.sin(np.pi / ny) ** 2)
    plan = halhip.Diffusion2D(ctx, nx, ny)
    plan.upload(f)
    for k in (1, 2):
        plan.run_steps(n, r, 0.0, True, True)
        np.testing.assert_allclose(plan.download(), lam ** (k * n) * f, rtol=0, atol=1e-6)


@pytest.mark.parametrize("wrap", [True, False])
def test_resident_mass_conservation(ctx, wrap):
    # tests/PDEGridTests.cpp:72-104: periodic and zero-flux diffusion keep
    # the sum of the field.
    X, Y = 48, 40
    f = np.random.default_rng(11).random(X * Y).astype(np.float32)
    plan = halhip.Diffusion2D(ctx, X, Y)
    plan.upload(f)
    plan.run_steps(200, 0.2, 0.0, wrap, wrap)
    got = plan.download().astype(np.float64).sum()
    m0 = f.astype(np.float64).sum()
    assert abs(got - m0) < 1e-5 * m0


def test_uniform_per_cell_rates_equal_uniform_stencil(ctx):
    # tests/PDEGridTests.cpp:291-330: per-cell rates that are all equal give
    # the constant-rate stencil.
    X, Y, r = 13, 17, 0.15
    f = np.random.default_rng(12).random(X * Y)
    plan = halhip.Diffusion2D(ctx, X, Y)
    for wx in (0, 1):
        for wy in (0, 1):
            want = plan.diffuse(f, np.zeros(X * Y), r, wx, wy)
            got = plan.diffuse_rates(f, np.zeros(X * Y), np.full(X * Y, r), np.full(X * Y, r), wx, wy)
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-7)


@pytest.mark.parametrize("X,Y", [(1000, 704), (1536, 1088)])
def test_per_call_in_pieces_matches_oracle(ctx, oracle, X, Y):
    # Per-call arrays of 256K values and more are copied in several pieces,
    # the host converting one piece while another is copied: 704,000 and
    # 1,671,168 cells make 2 and 6 pieces of uneven size.  The deltas start
    # non-zero, so the accumulation is checked too.
    f = np.random.default_rng(X + Y).random(X * Y)
    d0 = np.random.default_rng(1).random(X * Y)
    for adi in (False, True):
        plan = halhip.Adi2D(ctx, X, Y) if adi else halhip.Diffusion2D(ctx, X, Y)
        rate = 0.8 if adi else 0.2
        got = plan.diffuse(f, d0.copy(), rate, True, False)
        want = (oracle.adi if adi else oracle.diffuse)(f, d0.copy(), rate, X, Y, 1, 0)
        np.testing.assert_allclose(got, want, rtol=0, atol=3e-6, err_msg=f"adi={adi}")
        plan.close()


def test_rates_interface(ctx, golden, oracle):
    f, rx, ry = golden["rates_field"], golden["rates_x"], golden["rates_y"]
    plan = halhip.Diffusion2D(ctx, 9, 11)
    for wx in (0, 1):
        for wy in (0, 1):
            got = plan.diffuse_rates(f, np.zeros(99), rx, ry, wx, wy)
            np.testing.assert_allclose(got, golden[f"rates_{wx}{wy}_out"], atol=2e-6)


def test_size_mismatch_raises(ctx):
    # tests/IDiffusionBackend2DTests.cpp:146-177
    plan = halhip.Diffusion2D(ctx, 3, 4)
    with pytest.raises(halhip.InvalidArgument):
        plan.diffuse(np.zeros(12), np.zeros(11), 0.1, False, False)


@pytest.mark.parametrize("wrap", [(1, 1), (0, 0), (1, 0)])
def test_resident_ticks_match_oracle(ctx, oracle, wrap):
    X, Y = 37, 70
    rng = np.random.default_rng(11)
    f = rng.random(X * Y).astype(np.float32)
    plan = halhip.Diffusion2D(ctx, X, Y)
    plan.upload(f)
    plan.set_sources([5, 5, 1000], [1.0, 0.5, 2.0])
    plan.run_steps(120, 0.2, 2e-3, *wrap)
    got = plan.download()
    ref = oracle.explicit_ticks(f.astype(np.float64), 120, 0.2, 2e-3, X, Y, *wrap, [5, 5, 1000], [1.0, 0.5, 2.0])
    assert rel_l2(got, ref) < 1e-6


def test_dirichlet_resident_matches_oracle(ctx, oracle):
    X, Y = 40, 33
    f = np.random.default_rng(2).random(X * Y).astype(np.float32)
    plan = halhip.Diffusion2D(ctx, X, Y)
    plan.upload(f)
    plan.run_steps(30, 0.25, 0.0, False, False, bc=0.5)
    got = plan.download()
    ref = f.astype(np.float64)
    for _ in range(30):
        ref = ref + oracle.diffuse(ref, np.zeros_like(ref), 0.25, X, Y, 0, 0, 0.5)
    assert rel_l2(got, ref) < 1e-6


def test_workload_256_1000_ticks_within_1e5(ctx, oracle):
    w = workloads.explicit_256()
    plan = halhip.Diffusion2D(ctx, w["X"], w["Y"])
    plan.upload(w["field"])
    plan.set_sources(w["src_idx"], w["src_val"])
    plan.run_steps(1000, w["rate"], w["decay"], w["wrapX"], w["wrapY"])
    got = plan.download()
    ref = oracle.explicit_ticks(w["field"].astype(np.float64), 1000, w["rate"], w["decay"], w["X"], w["Y"],
                                1, 1, w["src_idx"], w["src_val"].astype(np.float64))
    err = rel_l2(got, ref)
    print(f"explicit 256^2 x 1000 ticks: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-5


def _explicit_ref(oracle, f, ticks, rate, decay, X, Y, wx, wy, bc, idx, val):
    """FP64 tick composition Diffusion -> MulAll(-k) -> Add -> Update with a
    Dirichlet boundary (the oracle's explicit_ticks covers bc=None)."""
    if bc is None:
        return oracle.explicit_ticks(f.astype(np.float64), ticks, rate, decay, X, Y, wx, wy, idx,
                                     np.asarray(val, np.float64))
    u = f.astype(np.float64)
    for _ in range(ticks):
        d = oracle.diffuse(u, np.zeros_like(u), rate, X, Y, wx, wy, bc)
        d -= decay * u
        np.add.at(d, np.asarray(idx, np.int64), np.asarray(val, np.float64))
        u = u + d
    return u


# Several ticks per launch (yDim <= 1024) and the single-tick kernel (longer
# rows): tiny periodic grids whose halo wraps more than once, partial last
# workgroups, remainder launches, a dense source map (> 64 sources) and
# Dirichlet edges.
MULTI_CASES = [
    (3, 50, 1, 1, None, 3, 37),
    (5, 64, 0, 0, 0.3, 0, 13),
    (70, 257, 1, 0, None, 100, 50),
    (33, 1024, 0, 1, 0.7, 2, 9),
    (9, 1500, 1, 1, None, 2, 11),
    (300, 17, 0, 0, 0.0, 5, 101),
    (1, 1, 1, 1, None, 1, 20),
]


@pytest.mark.parametrize("X,Y,wx,wy,bc,nsrc,ticks", MULTI_CASES)
def test_multi_tick_launches_match_oracle(ctx, oracle, X, Y, wx, wy, bc, nsrc, ticks):
    rng = np.random.default_rng(X * 31 + Y)
    f = rng.random(X * Y).astype(np.float32)
    idx = rng.integers(0, X * Y, nsrc)
    val = rng.random(nsrc).astype(np.float32)
    plan = halhip.Diffusion2D(ctx, X, Y)
    plan.upload(f)
    if nsrc:
        plan.set_sources(idx, val)
    plan.run_steps(ticks, 0.2, 1e-3, wx, wy, bc=bc)
    got = plan.download()
    ref = _explicit_ref(oracle, f, ticks, 0.2, 1e-3, X, Y, wx, wy, bc, idx, val)
    assert rel_l2(got, ref) < 1e-6, (X, Y, rel_l2(got, ref))


def test_multi_tick_shapes_are_bit_identical(ctx, oracle):
    # Every (rows per workgroup, ticks per launch) gives the same bits as one
    # launch per tick, and matches the FP64 oracle.
    X, Y, ticks = 40, 256, 37
    f = np.random.default_rng(12).random(X * Y).astype(np.float32)
    idx = np.array([5 * Y + 9, 39 * Y + 255], np.int64)
    val = np.array([1.0, 0.5], np.float32)
    outs = {}
    for shape in ((0, 0), (1, 8), (2, 15), (8, 8), (4, 12), (1, 1)):
        p = halhip.Diffusion2D(ctx, X, Y)
        p.set_multi_shape(*shape)
        if shape != (0, 0):
            assert p.multi_shape() == shape
        p.set_sources(idx, val)
        p.upload(f)
        p.run_steps(ticks, 0.2, 1e-3, 1, 0, bc=None)
        outs[shape] = p.download()
        p.close()
    for shape, got in outs.items():
        np.testing.assert_array_equal(got, outs[(0, 0)], err_msg=str(shape))
    ref = oracle.explicit_ticks(f.astype(np.float64), ticks, 0.2, 1e-3, X, Y, 1, 0, src_idx=idx,
                                src_val=val.astype(np.float64))
    assert rel_l2(outs[(0, 0)], ref) < 1e-5
    p = halhip.Diffusion2D(ctx, X, Y)
    with pytest.raises(halhip.InvalidArgument):
        p.set_multi_shape(4, 40)      # 84 rows held: more than fit
    p.close()
