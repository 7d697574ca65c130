"""Pins the FP64 oracle (oracle/hal_oracle.cpp) before it is trusted.

Each check is independent of the oracle's own formulas: golden vectors from
the compiled reference, closed forms the reference's tests use, textbook
solvers (dense / banded), brute force and conservation laws.
"""
import numpy as np
import pytest
import scipy.linalg

from helpers import eigen_field, golden_cases, lambda_cn


# ---- golden vectors from the compiled reference ------------------------------

def test_rng_matches_reference_stream(oracle, golden):
    assert np.array_equal(oracle.rng_doubles(42, 64), golden["rng42_doubles"])
    assert np.array_equal(oracle.rng_u64(1234, 16), golden["rng1234_u64"])
    assert np.array_equal(oracle.rng_doubles(7, 16), golden["rng7_doubles"])


def test_rng_known_answer():
    # splitmix64 from seed 0 yields 0xE220A8397B1DCDAF first (Vigna's
    # published reference output); xoshiro256** state = 4 splitmix outputs.
    from oracle import Oracle
    o = Oracle()
    sm = 0
    state = []
    for _ in range(4):
        sm = (sm + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = sm
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        state.append(z ^ (z >> 31))
    assert state[0] == 0xE220A8397B1DCDAF
    s1 = state[1]
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & (2**64 - 1)
    first = (rotl((s1 * 5) & (2**64 - 1), 7) * 9) & (2**64 - 1)
    assert int(o.rng_u64(0, 1)[0]) == first


def test_diffusion_and_adi_match_golden(oracle, golden):
    for c in golden_cases(golden):
        fn = oracle.adi if c["adi"] else oracle.diffuse
        got = fn(c["field"], c["deltas_in"], c["rate"], c["X"], c["Y"], c["wrapX"], c["wrapY"], c["bc"])
        np.testing.assert_array_equal(got, c["deltas_out"], err_msg=f"case {c['k']}")


def test_rates_match_golden(oracle, golden):
    f, rx, ry = golden["rates_field"], golden["rates_x"], golden["rates_y"]
    for wx in (0, 1):
        for wy in (0, 1):
            got = oracle.diffuse_rates(f, np.zeros(99), rx, ry, 9, 11, wx, wy)
            np.testing.assert_array_equal(got, golden[f"rates_{wx}{wy}_out"])


def test_tick_compositions_match_golden(oracle, golden):
    got = oracle.explicit_ticks(golden["ticks_explicit_field"], 25, 0.2, 1e-3, 16, 16, 1, 1, [8 * 16 + 8], [1.0])
    np.testing.assert_array_equal(got, golden["ticks_explicit_out"])
    got = oracle.adi_ticks(golden["ticks_adi_field"], 6, 5.0, 5e-3, 24, 20, 0, 0, bc=0.25,
                           src_idx=[3, 77, 77, 400], src_val=[1.0, 0.5, 0.25, 2.0])
    np.testing.assert_array_equal(got, golden["ticks_adi_out"])
    got = oracle.adi_ticks(golden["ticks_adi_field"], 6, 0.8, 0.0, 24, 20, 0, 1)
    np.testing.assert_array_equal(got, golden["ticks_adi_zf_out"])


def test_agent_forces_match_golden(oracle, golden):
    pos, rad = golden["agents_pos"], golden["agents_rad"]
    np.testing.assert_array_equal(oracle.agent_forces(pos, rad, 3.0, 0.4, 16, 16, 1, 1), golden["agents_wrap_out"])
    np.testing.assert_array_equal(oracle.agent_forces(pos, rad, 3.0, 0.4, 16, 16, 0, 0), golden["agents_nowrap_out"])
    np.testing.assert_array_equal(oracle.agent_forces(golden["agents2_pos"], np.full(500, 0.5), 1.0, 1.0, 40, 24, 0, 1),
                                  golden["agents2_out"])


# ---- closed forms (the reference's own known answers) -----------------------

def test_explicit_periodic_eigenmode(oracle):
    # tests/IDiffusionBackend2DTests.cpp:23-72
    nx, ny, r = 16, 20, 0.05
    f = eigen_field(nx, ny)
    got = f + oracle.diffuse(f, np.zeros_like(f), r, nx, ny, 1, 1)
    lam = 1 - 4 * r * (np.sin(np.pi / nx) ** 2 + np.sin(np.pi / ny) ** 2)
    np.testing.assert_allclose(got, lam * f, rtol=0, atol=1e-12)


@pytest.mark.parametrize("r", [0.3, 5.0])
def test_adi_periodic_eigenmode_is_crank_nicolson_product(oracle, r):
    # tests/IDiffusionBackend2DTests.cpp:76-144, tests/PDEADITests.cpp:75-130
    nx, ny = 12, 18
    f = eigen_field(nx, ny)
    got = f + oracle.adi(f, np.zeros_like(f), r, nx, ny, 1, 1)
    lam = lambda_cn(r, nx) * lambda_cn(r, ny)
    np.testing.assert_allclose(got, lam * f, rtol=0, atol=1e-12)


@pytest.mark.parametrize("wrap", [(0, 0), (1, 0), (0, 1)])
def test_mass_conservation_zero_flux(oracle, wrap):
    # tests/PDEGridTests.cpp:72-104 (explicit), tests/AdiKernel2DTests.cpp:70-141 (ADI)
    rng = np.random.default_rng(3)
    X, Y = 20, 14
    f = rng.random(X * Y)
    for fn, rate in ((oracle.diffuse, 0.2), (oracle.adi, 0.6), (oracle.adi, 5.0)):
        d = fn(f, np.zeros_like(f), rate, X, Y, *wrap)
        assert abs(d.sum()) < 1e-11


def test_decay_only_closed_form(oracle):
    # u = c everywhere, no diffusion flux: u_t = c (1-k)^t exactly.
    f = np.full(64, 3.0)
    got = oracle.explicit_ticks(f, 50, 0.2, 1e-2, 8, 8, 1, 1)
    np.testing.assert_allclose(got, 3.0 * (1 - 1e-2) ** 50, rtol=1e-13)


def test_adi_dirichlet_fixed_point_and_relaxation(oracle):
    # tests/PDEADITests.cpp:219-264
    f = np.full(7 * 5, 4.0)
    d = oracle.adi(f, np.zeros_like(f), 0.9, 7, 5, 0, 0, 4.0)
    np.testing.assert_allclose(d, 0.0, atol=1e-13)
    got = oracle.adi_ticks(np.zeros(8 * 6), 400, 1.5, 0.0, 8, 6, 1, 0, bc=2.0)
    np.testing.assert_allclose(got, 2.0, rtol=1e-6)


def test_adi_dirichlet_linear_ramp_is_steady(oracle):
    # tests/PDEADITests.cpp:297-354: bc 1 at x=-1, 3 at x=X, ramp between faces.
    X, Y = 6, 4
    ramp = lambda x: 1.0 if x < 0 else (3.0 if x >= X else 1.0 + 2.0 * (x + 0.5) / X)
    f = np.array([ramp(x) for x in range(X) for _ in range(Y)])
    bc = (np.full(Y, 1.0), np.full(Y, 3.0), np.array([ramp(x) for x in range(X)]), np.array([ramp(x) for x in range(X)]))
    d = oracle.adi(f, np.zeros_like(f), 0.9, X, Y, 0, 0, bc)
    np.testing.assert_allclose(d, 0.0, atol=1e-12)


# ---- textbook solvers ------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 7, 33])
def test_thomas_vs_dense_solve(oracle, n):
    rng = np.random.default_rng(n)
    a, c = -rng.random(n), -rng.random(n)
    b = 3 + rng.random(n)
    d = rng.random(n)
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    np.testing.assert_allclose(oracle.thomas(a, b, c, d), np.linalg.solve(M, d), rtol=1e-12, atol=1e-14)
    if n >= 3:
        # TDMAperiodic (TdmaSolver.h:68-120) takes alpha = a[0] and beta =
        # c[n-1] and pairs them so that the solved cyclic matrix has top-right
        # corner c[n-1] and bottom-left a[0]; HAL only calls it with equal
        # off-diagonals, where the pairing does not matter.
        P = M.copy()
        P[0, n - 1] = c[n - 1]
        P[n - 1, 0] = a[0]
        np.testing.assert_allclose(oracle.thomas(a, b, c, d, periodic=True), np.linalg.solve(P, d),
                                   rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("n,rate,has_bc", [(5, 5.0, False), (64, 5.0, True), (4096, 5.0, True), (100, 0.3, False)])
def test_adi_line_vs_banded_solve(oracle, n, rate, has_bc):
    # ADISolveRow coefficients (ADIequations.h:225-270) assembled by hand.
    rng = np.random.default_rng(n)
    d = rng.random(n)
    off = -rate / 2
    diag = np.full(n, 1 + rate)
    adj = rate / 2 if has_bc else -rate / 2
    diag[0] += adj
    diag[-1] += adj
    ab = np.zeros((3, n))
    ab[0, 1:] = off
    ab[1] = diag
    ab[2, :-1] = off
    want = scipy.linalg.solve_banded((1, 1), ab, d)
    np.testing.assert_allclose(oracle.adi_line(d, rate, 0, has_bc), want, rtol=1e-12, atol=1e-14)


# ---- agents: brute force and sorted bins ----------------------------------

def _brute(pos, rad, R, k, X, Y, wx, wy):
    n = rad.size
    out = np.zeros(2 * n)
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            d = []
            for ax, dim, w in ((0, X, wx), (1, Y, wy)):
                a, b = pos[2 * i + ax], pos[2 * j + ax]
                v = b - a
                if w and abs(v) > dim * 0.5:
                    v = v + dim if a > b else v - dim
                d.append(v)
            dist = np.hypot(*d)
            if 0 < dist < R:
                F = k * ((rad[i] + rad[j]) - dist)
                out[2 * i] -= d[0] / dist * F
                out[2 * i + 1] -= d[1] / dist * F
    return out


def test_agent_forces_vs_brute_force(oracle, golden):
    # tests/IAgentForceBackend2DTests.cpp:51-121
    pos, rad = golden["agents_pos"], golden["agents_rad"]
    for wx, wy in ((1, 1), (0, 0), (1, 0)):
        np.testing.assert_allclose(oracle.agent_forces(pos, rad, 3.0, 0.4, 16, 16, wx, wy),
                                   _brute(pos, rad, 3.0, 0.4, 16, 16, wx, wy), rtol=1e-12, atol=1e-12)


def test_coincident_agents_give_zero(oracle):
    # tests/IAgentForceBackend2DTests.cpp:153-165
    v = oracle.agent_forces([4.0, 4.0, 4.0, 4.0], [0.5, 0.5], 2.0, 1.0, 8, 8, 0, 0)
    assert np.all(v == 0.0)


def test_agent_bins_are_stable_counting_sort(oracle):
    rng = np.random.default_rng(5)
    X, Y, n = 13, 7, 400
    pos = rng.random(2 * n) * np.tile([X + 2.0, Y + 2.0], n) - 1.0   # some off-grid (clamped)
    offsets, bucket = oracle.agent_bins(pos, X, Y)
    cx = np.clip(np.floor(pos[0::2]).astype(int), 0, X - 1)
    cy = np.clip(np.floor(pos[1::2]).astype(int), 0, Y - 1)
    cell = cx * Y + cy
    order = np.lexsort((np.arange(n), cell))
    np.testing.assert_array_equal(bucket, order)
    np.testing.assert_array_equal(offsets, np.concatenate([[0], np.cumsum(np.bincount(cell, minlength=X * Y))]))


def test_agent_move_safe_semantics(oracle):
    # AgentPT2D::MoveSafePT (include/Hal/Agents/AgentPT2D.h:157-186)
    pos = np.array([1.0, 1.0, 9.5, 2.0, 5.0, 0.05])
    vel = np.array([1.0, 1.0, 10.0, 1.0, 0.0, -1.0])
    got = oracle.agent_move(pos, vel, 0.1, 10, 10, 0, 0)
    np.testing.assert_allclose(got, [1.1, 1.1, 9.5, 2.1, 5.0, 0.05])
    got = oracle.agent_move(pos, vel, 0.1, 10, 10, 1, 1)
    np.testing.assert_allclose(got, [1.1, 1.1, 0.5, 2.1, 5.0, 9.95])


def test_agent_motion_off_equals_plain_ticks(oracle):
    rng = np.random.default_rng(21)
    X, Y, n = 12, 10, 150
    pos = rng.random(2 * n) * np.tile([X, Y], n)
    rad = np.full(n, 0.5)
    got, carry = oracle.agent_ticks_motion(pos, np.zeros(2 * n), rad, 5, 1.0, 1.0, 0.1, 0.0, 0.0, X, Y, 1, 0)
    np.testing.assert_array_equal(got, oracle.agent_ticks(pos, rad, 5, 1.0, 1.0, 0.1, X, Y, 1, 0))
    assert np.all(carry == 0.0)


def test_agent_motion_friction_and_cap_closed_forms(oracle):
    # tests/SphericalAgentTests.cpp:88-102: ApplyFriction(0.5) on (3, 4) gives
    # (1.5, 2); CapVelocity(2) on (3, 4) gives norm 2.  A lone agent feels no
    # force, so its tick is exactly carry -> cap -> move -> friction.
    pos = np.array([10.0, 10.0])
    rad = np.array([0.5])
    p1, c1 = oracle.agent_ticks_motion(pos, [3.0, 4.0], rad, 1, 1.0, 1.0, 1.0, 0.5, 0.0, 40, 40, 0, 0)
    np.testing.assert_allclose(c1, [1.5, 2.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(p1, [13.0, 14.0], rtol=0, atol=1e-13)
    # cap before the move, friction after it
    p2, c2 = oracle.agent_ticks_motion(pos, [3.0, 4.0], rad, 1, 1.0, 1.0, 1.0, 0.5, 2.0, 40, 40, 0, 0)
    np.testing.assert_allclose(p2, [11.2, 11.6], rtol=0, atol=1e-13)
    np.testing.assert_allclose(c2, [0.6, 0.8], rtol=0, atol=1e-15)
    # geometric series over t ticks: moved dt*v0*(1-f^t)/(1-f), carry f^t*v0
    f, dt, t = 0.8, 0.25, 7
    p3, c3 = oracle.agent_ticks_motion(pos, [2.0, -1.0], rad, t, 1.0, 1.0, dt, f, 0.0, 40, 40, 0, 0)
    s = (1 - f ** t) / (1 - f)
    np.testing.assert_allclose(p3, [10.0 + dt * 2.0 * s, 10.0 - dt * 1.0 * s], rtol=0, atol=1e-12)
    np.testing.assert_allclose(c3, [2.0 * f ** t, -1.0 * f ** t], rtol=1e-13, atol=0)


def test_agent_motion_adds_forces_to_carried_velocity(oracle):
    # SumForces accumulates into the carried velocity (SphericalAgent2D.h:100-128).
    pos = np.array([5.0, 5.0, 5.6, 5.1])
    rad = np.array([0.5, 0.5])
    v0 = np.array([0.3, -0.2, 0.0, 0.4])
    F = oracle.agent_forces(pos, rad, 1.0, 1.0, 16, 16, 0, 0)
    assert np.abs(F).max() > 0.1
    p1, c1 = oracle.agent_ticks_motion(pos, v0, rad, 1, 1.0, 1.0, 0.5, 1.0, 0.0, 16, 16, 0, 0)
    np.testing.assert_allclose(c1, F + v0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p1, pos + 0.5 * (F + v0), rtol=0, atol=1e-14)


# ---- per-cell-diffusivity ADI and adjoint (no reference counterpart) ------

def test_cell_adi_uniform_equals_reference_adi(oracle):
    # Uniform D reduces to the reference's zero-flux ADI tick + MulAll(-k).
    rng = np.random.default_rng(11)
    X, Y = 13, 9
    u0 = rng.random(X * Y)
    got = oracle.adi_cell_forward(u0, np.full(X * Y, 2.5), 3e-3, 4, X, Y)
    ref = oracle.adi_ticks(u0, 4, 2.5, 3e-3, X, Y, 0, 0)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-14)


def test_cell_adi_tick_transpose_dot_product(oracle):
    # <M x, y> = <x, M^T y> for the tick's linear map M at fixed D.
    rng = np.random.default_rng(12)
    X, Y = 11, 7
    D = 0.5 + rng.random(X * Y)
    x, y = rng.random(X * Y), rng.random(X * Y)
    Mx = oracle.adi_cell_forward(x, D, 0.02, 1, X, Y)
    MTy = oracle.adi_cell_tick_transpose(y, D, 0.02, X, Y)
    assert abs(Mx @ y - x @ MTy) < 1e-12 * abs(Mx @ y)


def test_cell_adi_gradient_vs_central_differences(oracle):
    rng = np.random.default_rng(13)
    X, Y, ticks, k = 10, 8, 12, 5e-3
    D = 0.5 + rng.random(X * Y)
    u0 = rng.random(X * Y)
    target = oracle.adi_cell_forward(u0, D * (1 + 0.1 * rng.random(X * Y)), k, ticks, X, Y)
    J, gD, gk = oracle.adi_cell_gradient(u0, D, k, ticks, X, Y, target)
    h = 1e-6
    def loss(Dv, kv):
        uT = oracle.adi_cell_forward(u0, Dv, kv, ticks, X, Y)
        return 0.5 * np.sum((uT - target) ** 2)
    assert abs(J - loss(D, k)) < 1e-14
    for c in rng.choice(X * Y, 12, replace=False):
        Dp, Dm = D.copy(), D.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (loss(Dp, k) - loss(Dm, k)) / (2 * h)
        assert abs(fd - gD[c]) < 1e-6 * max(1.0, abs(fd)), (c, fd, gD[c])
    fdk = (loss(D, k + h) - loss(D, k - h)) / (2 * h)
    assert abs(fdk - gk) < 1e-6 * max(1.0, abs(fdk))


# Boundary kinds of the per-cell model: (wrapX, wrapY, bc) with bc None (zero
# flux), a constant Dirichlet value, or "edges" (seeded edge arrays).
CELL_BOUNDS = [(1, 1, None), (1, 0, None), (0, 1, None), (0, 0, 0.7), (1, 0, 0.4), (0, 1, -0.3),
               (0, 0, "edges"), (1, 0, "edges")]


def cell_bc(bc, X, Y, seed=5):
    if isinstance(bc, str):
        rng = np.random.default_rng(seed)
        return (rng.random(Y), rng.random(Y), rng.random(X), rng.random(X))
    return bc


@pytest.mark.parametrize("wx,wy,bc", CELL_BOUNDS)
def test_cell_adi_uniform_equals_reference_adi_every_boundary(oracle, wx, wy, bc):
    # Uniform D reduces to the reference's ADI tick (cyclic lines, Dirichlet
    # diagonal 1 + 3r/2 and right-hand side r*bc) + MulAll(-k).
    rng = np.random.default_rng(21)
    X, Y = 13, 9
    b = cell_bc(bc, X, Y)
    u0 = rng.random(X * Y)
    got = oracle.adi_cell_forward(u0, np.full(X * Y, 2.5), 3e-3, 4, X, Y, wx, wy, b)
    ref = oracle.adi_ticks(u0, 4, 2.5, 3e-3, X, Y, wx, wy, bc=b)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("wx,wy", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_cell_explicit_half_step_is_the_per_cell_rate_stencil(oracle, wx, wy):
    # (I + L/2) p - p along one axis is half the reference's per-cell-rate
    # stencil (src/Pde/PDEequations.cpp:199-228) with that axis's rates D and
    # the other axis's 0: the same face convention, wrap face included.
    rng = np.random.default_rng(22)
    X, Y = 7, 6
    D = 0.3 + rng.random(X * Y)
    p = rng.random(X * Y)
    zero = np.zeros(X * Y)
    for axis in (0, 1):
        got = oracle.adi_cell_explicit(p, D, X, Y, axis, wx, wy) - p
        rx, ry = (D, zero) if axis == 0 else (zero, D)
        want = 0.5 * oracle.diffuse_rates(p, zero, rx, ry, X, Y, wx, wy)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("wx,wy,bc", CELL_BOUNDS)
def test_cell_adi_tick_transpose_dot_product_every_boundary(oracle, wx, wy, bc):
    # <M x, y> = <x, M^T y> for the linear part M x = F(x) - F(0) of the tick.
    rng = np.random.default_rng(23)
    X, Y = 11, 7
    b = cell_bc(bc, X, Y)
    D = 0.5 + rng.random(X * Y)
    x, y = rng.random(X * Y), rng.random(X * Y)
    F0 = oracle.adi_cell_forward(np.zeros(X * Y), D, 0.02, 1, X, Y, wx, wy, b)
    Mx = oracle.adi_cell_forward(x, D, 0.02, 1, X, Y, wx, wy, b) - F0
    MTy = oracle.adi_cell_tick_transpose(y, D, 0.02, X, Y, wx, wy, b)
    assert abs(Mx @ y - x @ MTy) < 1e-12 * abs(Mx @ y)


@pytest.mark.parametrize("wx,wy,bc", CELL_BOUNDS)
def test_cell_adi_gradient_vs_central_differences_every_boundary(oracle, wx, wy, bc):
    rng = np.random.default_rng(24)
    X, Y, ticks, k = 9, 8, 10, 5e-3
    b = cell_bc(bc, X, Y)
    D = 0.5 + rng.random(X * Y)
    u0 = rng.random(X * Y)
    target = oracle.adi_cell_forward(u0, D * (1 + 0.1 * rng.random(X * Y)), k, ticks, X, Y, wx, wy, b)
    J, gD, gk = oracle.adi_cell_gradient(u0, D, k, ticks, X, Y, target, wx, wy, b)
    h = 1e-6

    def loss(Dv, kv):
        uT = oracle.adi_cell_forward(u0, Dv, kv, ticks, X, Y, wx, wy, b)
        return 0.5 * np.sum((uT - target) ** 2)

    assert abs(J - loss(D, k)) < 1e-14
    # every corner, a cell of every edge, and interior cells
    cells = [0, Y - 1, (X - 1) * Y, X * Y - 1, 3, 3 * Y, 4 * Y + Y - 1, (X - 1) * Y + 2]
    cells += list(rng.choice(X * Y, 6, replace=False))
    for c in cells:
        Dp, Dm = D.copy(), D.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (loss(Dp, k) - loss(Dm, k)) / (2 * h)
        assert abs(fd - gD[c]) < 1e-6 * max(1.0, abs(fd)), (c, fd, gD[c])
    fdk = (loss(D, k + h) - loss(D, k - h)) / (2 * h)
    assert abs(fdk - gk) < 1e-6 * max(1.0, abs(fdk))


def test_cell_adi_periodic_two_cell_lines(oracle):
    # A wrapped axis of two cells has two faces between the same cells; the
    # tick must still match the reference's cyclic solve for uniform D.
    rng = np.random.default_rng(25)
    X, Y = 2, 5
    u0 = rng.random(X * Y)
    got = oracle.adi_cell_forward(u0, np.full(X * Y, 1.5), 0.0, 3, X, Y, 1, 1)
    ref = oracle.adi_ticks(u0, 3, 1.5, 0.0, X, Y, 1, 1)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-14)


def test_cell_adi_gradient_at_workload_size_vs_fixture(oracle):
    """The FP64 adjoint on the 1024^2, 200-tick adjoint workload against the
    committed FP64 central differences (tests/golden/make_adjoint_fd.py)."""
    import os
    import workloads
    fd = np.load(os.path.join(os.path.dirname(__file__), "golden", "adjoint_fd_1024.npz"), allow_pickle=False)
    w = workloads.adjoint_problem()
    X, Y, ticks = w["X"], w["Y"], w["ticks"]
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)
    k = float(np.float32(w["decay"]))
    u0 = w["u0"].astype(np.float64)
    target = f32(oracle.adi_cell_forward(u0, f32(w["D_target"]), k, ticks, X, Y))
    J, gD, gk = oracle.adi_cell_gradient(u0, f32(w["D"]), k, ticks, X, Y, target)
    assert abs(J - fd["loss"][0]) < 1e-12 * J
    scale = np.abs(gD).max()
    g = gD[fd["cells"]]
    # central differences at h = 1e-6 carry about eps J / h (~1.3e-12) of
    # rounding in the loss difference; the bound allows 8 times that
    atol = 8 * np.finfo(np.float64).eps * J / fd["h"][0]
    np.testing.assert_allclose(g, fd["fd_gradD"], rtol=1e-5, atol=atol)
    assert np.abs(fd["fd_gradD"]).max() > 1e4 * atol  # some cells carry a real gradient
    assert scale >= np.abs(g).max()
    # dJ/dk: the fourth-order difference; the two-point one is off by its
    # h^2 J''' / 6 truncation (~1e-4 relative here: u_T is a degree-200
    # polynomial in k)
    assert abs(gk - fd["fd4_gradK"][0]) < 1e-7 * abs(gk)
    assert abs(gk - fd["fd_gradK"][0]) < 2e-4 * abs(gk)


# ---- upwind advection ----------------------------------------------------------

def test_advection_matches_golden_bit_for_bit(oracle, golden):
    from helpers import advection_cases
    n = 0
    for c in advection_cases(golden):
        got = oracle.advect(c["field"], c["deltas_in"], *c["vel"], c["X"], c["Y"], c["wrapX"], c["wrapY"], c["bc"])
        np.testing.assert_array_equal(got, c["uniform_out"], err_msg=str(c["k"]))
        got = oracle.advect_cells(c["field"], c["deltas_in"], c["xv"], c["yv"], c["X"], c["Y"], c["wrapX"],
                                  c["wrapY"], c["bc"], c["vbc"])
        np.testing.assert_array_equal(got, c["cells_out"], err_msg=str(c["k"]))
        n += 1
    assert n == 40


def test_model_ticks_match_golden(oracle, golden):
    f = golden["ticks_model_field"]
    got = oracle.model_ticks(f, 15, 0.15, 0.2, -0.1, 2e-3, 20, 18, 1, 0, None, [5, 200], [1.0, 0.5])
    np.testing.assert_array_equal(got, golden["ticks_model_out"])
    got = oracle.model_ticks(f, 15, 0.15, -0.05, 0.25, 0.0, 20, 18, 0, 0, 0.4, [7], [0.3])
    np.testing.assert_array_equal(got, golden["ticks_model_bc_out"])


def test_flux_form_conserves_mass_with_walls(oracle):
    # Zero-flux walls: every face flux leaves one cell and enters its
    # neighbour, so the deltas sum to zero; also on a periodic grid.
    rng = np.random.default_rng(3)
    X, Y = 13, 17
    f = rng.random(X * Y)
    xv, yv = rng.uniform(-0.3, 0.3, X * Y), rng.uniform(-0.3, 0.3, X * Y)
    for wx, wy in ((0, 0), (1, 1), (1, 0)):
        d = oracle.advect_cells(f, np.zeros(X * Y), xv, yv, X, Y, wx, wy)
        assert abs(d.sum()) < 1e-13 * np.abs(d).sum()


def test_uniform_velocity_field_matches_uniform_form(oracle):
    # A uniform velocity field reduces the flux form to u_upwind - u times
    # the velocity: equal to the uniform-velocity form everywhere on a
    # periodic grid and away from the walls otherwise.
    rng = np.random.default_rng(4)
    X, Y = 11, 9
    f = rng.random(X * Y)
    for vx, vy in ((0.2, -0.15), (-0.1, 0.3), (0.0, 0.25)):
        a = oracle.advect(f, np.zeros(X * Y), vx, vy, X, Y, 1, 1)
        b = oracle.advect_cells(f, np.zeros(X * Y), np.full(X * Y, vx), np.full(X * Y, vy), X, Y, 1, 1)
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-15)
        a = oracle.advect(f, np.zeros(X * Y), vx, vy, X, Y, 0, 0).reshape(X, Y)
        b = oracle.advect_cells(f, np.zeros(X * Y), np.full(X * Y, vx), np.full(X * Y, vy), X, Y, 0, 0).reshape(X, Y)
        np.testing.assert_allclose(b[1:-1, 1:-1], a[1:-1, 1:-1], rtol=0, atol=1e-15)


def test_upwind_shift_of_a_step(oracle):
    # Closed form: with velocity 1 along +x on a periodic grid one explicit
    # step moves every row to the next x exactly (u + (u[x-1] - u) * 1).
    rng = np.random.default_rng(5)
    X, Y = 8, 5
    f = rng.random(X * Y)
    d = oracle.advect(f, np.zeros(X * Y), 1.0, 0.0, X, Y, 1, 1)
    np.testing.assert_allclose((f + d).reshape(X, Y), np.roll(f.reshape(X, Y), 1, axis=0), rtol=0, atol=1e-15)
