"""Adjoint ADI on the GPU: forward vs the FP64 oracle, gradients vs the FP64
CPU adjoint and FP64 central finite differences, and a torch.autograd.Function
wrapper (test harness only) whose backward calls the C ABI adjoint."""
import numpy as np
import pytest

import halhip
import workloads
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def small_problem(X=24, Y=20, seed=0):
    rng = np.random.default_rng(seed)
    D = 0.5 + rng.random(X * Y)
    u0 = rng.random(X * Y).astype(np.float32)
    return D, u0


# Line lengths cover every per-thread width of the line solve (1, 2, 4, 8,
# 16 elements) with and without 16-byte vectors, including partly filled
# vector groups at the end of a line (2052 = 8 * 256 + 4).
@pytest.mark.parametrize("X,Y", [(24, 20), (64, 200), (130, 70), (300, 520), (1030, 18), (2052, 8)])
def test_forward_matches_oracle(ctx, oracle, X, Y):
    D, u0 = small_problem(X, Y, X)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, 5e-3, 30, 7)
    got = p.forward(u0)
    ref = oracle.adi_cell_forward(u0.astype(np.float64), D, 5e-3, 30, X, Y)
    assert rel_l2(got, ref) < 1e-5


def test_uniform_diffusivity_matches_plain_adi(ctx):
    X, Y = 50, 64
    u0 = np.random.default_rng(4).random(X * Y).astype(np.float32)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(np.full(X * Y, 2.0), 1e-3, 10, 10)
    plain = halhip.Adi2D(ctx, X, Y)
    plain.upload(u0)
    plain.run_steps(10, 2.0, 1e-3, False, False)
    assert rel_l2(p.forward(u0), plain.download()) < 1e-6


def f32(a):
    """The FP32 values the device holds, widened back to FP64."""
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def make_target(oracle, u0, D, k, ticks, X, Y, seed):
    # Generated with perturbed D and twice the decay: a target made with the
    # same k puts J(k) at the bottom of a steep parabola (curvature ~4e5 at
    # 40 ticks), where dJ/dk is a near-zero difference that rounding k to
    # FP32 alone moves by ~6e-4 relative.  Away from the minimum the FP32
    # rounding of the inputs moves dJ/dk by ~3e-8 relative.
    rng = np.random.default_rng(seed)
    return f32(oracle.adi_cell_forward(u0.astype(np.float64), D * (1 + 0.1 * rng.random(X * Y)), 2 * k, ticks, X, Y))


@pytest.mark.parametrize("every", [1, 6, 40])
def test_gradient_matches_fp64_adjoint(ctx, oracle, every):
    X, Y, ticks, k = 40, 36, 40, 5e-3
    D, u0 = small_problem(X, Y, 3)
    target = make_target(oracle, u0, D, k, ticks, X, Y, 9)
    J, gD, gk = oracle.adi_cell_gradient(u0.astype(np.float64), f32(D), float(np.float32(k)), ticks, X, Y, target)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, k, ticks, every)
    Jg, gDg, gkg = p.gradient(u0, target.astype(np.float32))
    assert abs(Jg - J) < 1e-4 * abs(J)
    assert rel_l2(gDg, gD) < 1e-4
    assert abs(gkg - gk) < 1e-4 * abs(gk)


def test_gradient_long_lines_match_fp64_adjoint(ctx, oracle):
    # x-lines of 516 cells: vector loads and four wavefronts per line in the
    # forward, recompute and adjoint solves
    X, Y, ticks, k = 516, 12, 10, 5e-3
    D, u0 = small_problem(X, Y, 4)
    target = make_target(oracle, u0, D, k, ticks, X, Y, 10)
    J, gD, gk = oracle.adi_cell_gradient(u0.astype(np.float64), f32(D), float(np.float32(k)), ticks, X, Y, target)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, k, ticks, 3)
    Jg, gDg, gkg = p.gradient(u0, target.astype(np.float32))
    assert abs(Jg - J) < 1e-4 * abs(J)
    assert rel_l2(gDg, gD) < 1e-4
    assert abs(gkg - gk) < 1e-4 * abs(gk)


def test_gradient_vs_central_differences(ctx, oracle):
    # FP64 central differences (h = 1e-6) of the FP64 model at seeded cells
    # and for the decay rate; the GPU gradient within 1e-4 (relative to the
    # gradient's scale).
    X, Y, ticks, k = 32, 28, 30, 5e-3
    D, u0 = small_problem(X, Y, 5)
    target = make_target(oracle, u0, D, k, ticks, X, Y, 6)
    rng = np.random.default_rng(7)
    p = halhip.Adjoint2D(ctx, X, Y)
    p.set_problem(D, k, ticks, 10)
    _, gD, gk = p.gradient(u0, target.astype(np.float32))
    u64 = u0.astype(np.float64)

    def loss(Dv, kv):
        return 0.5 * np.sum((oracle.adi_cell_forward(u64, Dv, kv, ticks, X, Y) - target) ** 2)

    h = 1e-6
    scale = np.abs(gD).max()
    worst = 0.0
    for c in rng.choice(X * Y, 64, replace=False):
        Dp, Dm = D.copy(), D.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (loss(Dp, k) - loss(Dm, k)) / (2 * h)
        worst = max(worst, abs(gD[c] - fd) / scale)
        assert abs(gD[c] - fd) < 1e-4 * scale, (c, gD[c], fd)
    fdk = (loss(D, k + h) - loss(D, k - h)) / (2 * h)
    print(f"adjoint vs FP64 central differences: D cells max |g - fd| / max|g| = {worst:.3e}, "
          f"decay |g - fd| / |fd| = {abs(gk - fdk) / abs(fdk):.3e}")
    assert abs(gk - fdk) < 1e-4 * abs(fdk)


def test_workload_1024_forward_first_ticks(ctx, oracle):
    w = workloads.adjoint_problem()
    p = halhip.Adjoint2D(ctx, w["X"], w["Y"])
    p.set_problem(w["D"], w["decay"], 3, 1)
    got = p.forward(w["u0"])
    ref = oracle.adi_cell_forward(w["u0"].astype(np.float64), w["D"], w["decay"], 3, w["X"], w["Y"])
    err = rel_l2(got, ref)
    print(f"adjoint workload 1024^2 forward, 3 ticks: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-5


def test_torch_autograd_function(ctx, oracle):
    torch = pytest.importorskip("torch")
    X, Y, ticks, k = 24, 20, 15, 5e-3
    D, u0 = small_problem(X, Y, 8)
    target = oracle.adi_cell_forward(u0.astype(np.float64), D * 1.05, k, ticks, X, Y).astype(np.float32)
    plan = halhip.Adjoint2D(ctx, X, Y)

    class Misfit(torch.autograd.Function):
        """J(D) = 1/2 |u_T(D) - target|^2 with the C ABI adjoint as backward."""

        @staticmethod
        def forward(fctx, Dt):
            plan.set_problem(Dt.detach().double().numpy(), k, ticks, 5)
            J, gD, _ = plan.gradient(u0, target)
            fctx.save_for_backward(torch.from_numpy(gD.astype(np.float64)))
            return Dt.new_tensor(J)

        @staticmethod
        def backward(fctx, gJ):
            (gD,) = fctx.saved_tensors
            return gJ * gD

    Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    J = Misfit.apply(Dt)
    J.backward()
    _, gD, _ = oracle.adi_cell_gradient(u0.astype(np.float64), D, k, ticks, X, Y, target.astype(np.float64))
    assert rel_l2(Dt.grad.numpy(), gD) < 1e-4
