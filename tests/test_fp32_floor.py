"""The FP32 floor of the reference ADI over long zero-flux runs with sources.

HAL's tick (Diffusion2DADI + MulAll + Add + Update, include/Hal/Pde/ADIequations.h:306-337,
TdmaSolver.h:27-66) restated with numpy float32 state and float32 Thomas
factors, on the smallest well of the 512-well workload (64 x 64, rate 4.30,
decay 6.7e-3, 32 sources), 500 ticks.  Against the FP64 oracle it lands at
~1.6e-5 relative L2, nearly all of it in the total mass, while the same
recurrences with FP64 factors and float32 state stay near 1e-7: the
rounding of the factors, not of the field, sets the floor that the GPU's
500-tick well results are held to (tests/test_workloads_gpu.py)."""
import numpy as np

import workloads
from helpers import rel_l2


def _ticks(spec, k, ticks, factor_dtype):
    X = int(spec["xdims"][k])
    r = float(spec["rate"][k])
    dec = float(spec["decay"][k])
    sel = spec["src_well"] == k
    off = -r / 2
    b = np.full(X, 1 + r)
    b[0] = b[-1] = 1 + r / 2
    bp = np.zeros(X)
    bp[0] = b[0]
    for i in range(1, X):
        bp[i] = b[i] - (off / bp[i - 1]) * off
    fw = np.zeros(X)
    fw[1:] = -(off / bp[:-1])
    bw = np.zeros(X)
    bw[:-1] = -off / bp[:-1]
    fw, ib, bw = (a.astype(factor_dtype) for a in (fw, 1 / bp, bw))
    f32 = np.float32

    def solve(d):  # lines along axis 1
        dp = d.copy()
        for i in range(1, X):
            dp[:, i] = fw[i] * dp[:, i - 1] + dp[:, i]
        x = np.empty_like(dp)
        x[:, -1] = ib[-1] * dp[:, -1]
        for i in range(X - 2, -1, -1):
            x[:, i] = bw[i] * x[:, i + 1] + ib[i] * dp[:, i]
        return x.astype(f32)

    def half(u, axis):  # u + r/2 * (neighbour differences), zero flux
        s = np.zeros_like(u)
        if axis == 1:
            s[:, :-1] += u[:, 1:] - u[:, :-1]
            s[:, 1:] += u[:, :-1] - u[:, 1:]
        else:
            s[:-1] += u[1:] - u[:-1]
            s[1:] += u[:-1] - u[1:]
        return (u + f32(r / 2) * s).astype(f32)

    u = np.zeros((X, X), f32)
    cells, vals = spec["src_cell"][sel], spec["src_val"][sel]
    for _ in range(ticks):
        us = solve(half(u, 1).T.copy()).T
        un = (solve(half(us, 0)) + f32(-dec) * u).astype(f32).ravel()
        for c, v in zip(cells, vals):
            un[c] += v
        u = un.reshape(X, X)
    return u.ravel()


def test_fp32_factors_set_the_long_run_floor(oracle):
    spec = workloads.wells_batch()
    k = int(np.argsort(spec["xdims"], kind="stable")[0])
    X = int(spec["xdims"][k])
    sel = spec["src_well"] == k
    ref = oracle.adi_ticks(np.zeros(X * X), 500, float(spec["rate"][k]), float(spec["decay"][k]), X, X, 0, 0,
                           src_idx=spec["src_cell"][sel], src_val=spec["src_val"][sel].astype(np.float64))
    f32_err = rel_l2(_ticks(spec, k, 500, np.float32), ref)
    f64_err = rel_l2(_ticks(spec, k, 500, np.float64), ref)
    print(f"64^2 well, 500 ticks vs FP64: float32 factors {f32_err:.3e}, float64 factors {f64_err:.3e}")
    assert 1.0e-5 < f32_err < 2e-5
    assert f64_err < 1e-6
