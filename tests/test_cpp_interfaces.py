"""The C++ drop-in boundary: Hal::Hip backends behind HAL's strategy
interfaces (include/Hal/Hip/HipBackends.h), exercised by model-side C++ code
(tests/cpp/hal_interface_tests.cpp)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
BIN = os.path.join(ROOT, "tests", "cpp", "hal_interface_tests")
LIB = os.path.join(ROOT, "hal-cpp_amd", "libhal_hip_backends.so")


def test_backends_library_exports_interface_classes():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for sym in ("Hal::Hip::HipDiffusionBackend2D::Diffuse", "Hal::Hip::HipAdiDiffusionBackend2D::DiffuseADI",
                "Hal::Hip::HipAgentForceBackend2D::AccumulateForces", "Hal::Internal::CheckFieldSizes",
                "Hal::Hip::HipContext::TryCreate", "Hal::Logger::Instance", "Hal::Hip::HipPDEGrid2D::Update",
                "Hal::Hip::HipPDEGrid2D::DiffusionADI"):
        assert sym in out, sym
    # vtables of the interface implementations are present
    assert re.search(r"vtable for Hal::Hip::HipAdiDiffusionBackend2D", out)


def test_cpp_program_skips_cleanly_without_gpu():
    import halhip
    if halhip.device_count() > 0:
        pytest.skip("host has a GPU")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 77, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_model_code_through_interfaces(tmp_path):
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def _draws(seed, n, scale, shift):
    import halhip
    return scale * (halhip.rng_doubles(seed, n) + shift)


def _edges(fn, X, Y):
    return (np.array([fn(-1, y) for y in range(Y)]), np.array([fn(X, y) for y in range(Y)]),
            np.array([fn(x, -1) for x in range(X)]), np.array([fn(x, Y) for x in range(X)]))


@pytest.mark.gpu
def test_cpp_model_code_on_the_device_grid_matches_the_oracle(tmp_path, oracle):
    """The device grid's scenarios in tests/cpp/hal_interface_tests.cpp
    (ResidentGridScenarios), replayed call by call on the FP64 oracle:
    fused ADI model ticks; a tick of every PDEGrid2D call kind with boundary
    callables; per-cell rates with uniform advection, SetAll(values) and
    fused explicit ticks (include/Hal/Grids/PDEGrid2D.h:89-399)."""
    from helpers import RefGrid, rel_l2
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    load = lambda name: np.fromfile(str(tmp_path / name), np.float64)

    # 1: fused ADI model ticks
    X, Y = 96, 128
    src = [0, 40 * Y + 77, X * Y - 1]
    ref = oracle.adi_ticks(_f32(_draws(11, X * Y, 1.0, 0.0)), 25, 5.0, 2e-3, X, Y, 0, 0, bc=0.0, src_idx=src,
                           src_val=[0.5, 1.0, 1.5])
    err = rel_l2(load("facade_adi.bin"), ref)
    print(f"device grid, fused ADI ticks vs FP64 oracle: rel L2 {err:.3e}")
    assert err < 1e-5

    # 2: every call kind in one tick, 8 ticks
    X, Y = 40, 36
    g = RefGrid(oracle, _f32(_draws(21, X * Y, 1.0, -0.2)), X, Y, (0, 0))
    xv = _draws(22, X * Y, 0.1, -0.5)
    yv = _draws(23, X * Y, 0.1, -0.5)
    bc = _edges(lambda x, y: 0.5 + 0.01 * x - 0.02 * y, X, Y)
    vbc = _edges(lambda x, y: 0.05 if (x < 0 or x >= X) else -0.03, X, Y)
    stats = []
    for t in range(8):
        g.diffuse(0.1, bc=bc)
        g.advect_cells(xv, yv, bc=bc, vbc=vbc)
        g.adi(2.0)
        g.point(0, 7 * Y + 5, 1.5)
        g.point(3, 100, 1.2)
        g.point(2, 200, -0.1)
        g.point(1, 300, 0.25)
        g.d += 1e-3
        g.d += g.u * -5e-3
        if t % 2 == 1:
            neg = g.u + g.d < 0
            g.d[neg] = -g.u[neg]
        stats += [np.abs(g.d).max(), np.abs(g.d / (np.abs(g.u) + 0.1)).max()]
        g.update()
    stats += [g.u.max(), g.u.min(), g.u.mean()]
    err = rel_l2(load("facade_general.bin"), g.u)
    got = load("facade_general_stats.bin")
    print(f"device grid, every call kind, 8 ticks vs FP64 oracle: rel L2 {err:.3e}, "
          f"MaxDelta / GetMax / GetMin / GetAvg max abs err {np.max(np.abs(got - stats)):.3e}")
    assert err < 1e-5
    np.testing.assert_allclose(got, stats, rtol=1e-5, atol=1e-6)

    # 3: per-cell rates + advection (periodic x), SetAll(values), fused ticks
    X, Y = 32, 40
    g = RefGrid(oracle, _f32(_draws(31, X * Y, 1.0, 0.0)), X, Y, (1, 0))
    rx = _draws(32, X * Y, 0.2, 0.0)
    ry = _draws(33, X * Y, 0.2, 0.0)
    for t in range(12):
        g.diffuse_rates(rx, ry)
        g.advect(0.1, 0.05)
        g.d += g.u * -1e-3
        g.point(1, 5, 0.1)
        g.update()
    g.d = _draws(34, X * Y, 1.0, 0.0) - g.u
    g.diffuse(0.2)
    g.update()
    g.diffuse(0.2)
    g.d += g.u * -1e-3
    g.update()
    err = rel_l2(load("facade_rates.bin"), g.u)
    print(f"device grid, per-cell rates + advection, SetAll, fused ticks vs FP64 oracle: rel L2 {err:.3e}")
    assert err < 1e-5
