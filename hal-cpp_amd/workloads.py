"""Seeded synthetic inputs of the benchmark workloads.

Every workload draws from Hal::Rng(seed).NextDouble() (src/Rng.cpp:91-96,
reproduced by halhip_rng_doubles) in a fixed order, evaluates in FP64 and
rounds to FP32 once; the FP64 checkers receive the same FP32 values widened
back, so both sides start from identical inputs.
"""
import numpy as np

import halhip


def explicit_256(seed=42, n=256):
    """Plain explicit diffusion: n x n field of U[0,1) in flat index order,
    point source 1.0 at the centre cell, added every tick (PDEGrid2D::Add)."""
    field = halhip.rng_doubles(seed, n * n).astype(np.float32)
    c = (n // 2) * n + (n // 2)
    return dict(X=n, Y=n, field=field, rate=0.2, decay=1e-3, wrapX=True, wrapY=True,
                src_idx=np.array([c], np.int64), src_val=np.array([1.0], np.float32))


def adi_blobs(n=4096, blobs=64, sigma=20.0, seed=42, ny=None):
    """ADI workload: sum of `blobs` Gaussian blobs on an n x ny grid (ny =
    n by default).  Per blob, in order from Rng(seed): cx = n*U, cy = ny*U,
    A = 0.5 + 0.5*U; value A*exp(-((x-cx)^2 + (y-cy)^2) / (2 sigma^2)) at
    integer cells, truncated at |dx|, |dy| <= 8 sigma.  Dirichlet boundary
    value 0, rate 5, 200 ticks."""
    ny = n if ny is None else ny
    u = halhip.rng_doubles(seed, 3 * blobs)
    field = np.zeros((n, ny), np.float64)
    half = int(round(8 * sigma))
    for b in range(blobs):
        cx, cy, amp = n * u[3 * b], ny * u[3 * b + 1], 0.5 + 0.5 * u[3 * b + 2]
        x0, x1 = max(0, int(np.floor(cx)) - half), min(n, int(np.floor(cx)) + half + 1)
        y0, y1 = max(0, int(np.floor(cy)) - half), min(ny, int(np.floor(cy)) + half + 1)
        xs = np.arange(x0, x1, dtype=np.float64)[:, None]
        ys = np.arange(y0, y1, dtype=np.float64)[None, :]
        field[x0:x1, y0:y1] += amp * np.exp(-((xs - cx) ** 2 + (ys - cy) ** 2) / (2 * sigma * sigma))
    return dict(X=n, Y=ny, field=field.astype(np.float32).ravel(), rate=5.0, bc=0.0, wrapX=False, wrapY=False)


def agents_uniform(n_agents=4_000_000, dim=2048, seed=42):
    """Agent workload: positions (dim*U, dim*U) per agent in order from
    Rng(seed), radius 0.5, cutoff 1.0, k = 1, dt = 0.1, no wrap."""
    u = halhip.rng_doubles(seed, 2 * n_agents)
    pos = (u * dim).astype(np.float32)
    rad = np.full(n_agents, 0.5, np.float32)
    return dict(X=dim, Y=dim, pos=pos, rad=rad, R=1.0, k=1.0, dt=0.1, wrapX=False, wrapY=False)


def wells_batch(n_wells=512, seed=42, lo=64, hi=1024, sources_per_well=32):
    """Batched wells: per well, in order from Rng(seed): side =
    round(exp(U(ln lo, ln hi))), D ~ U[0.1, 5], decay ~ U[0, 1e-2], then
    `sources_per_well` point sources (x = floor(side*U), y = floor(side*U),
    strength U[0,1]).  Square wells, zero flux, initial field 0."""
    per = 3 + 3 * sources_per_well
    u = halhip.rng_doubles(seed, n_wells * per).reshape(n_wells, per)
    side = np.rint(np.exp(np.log(lo) + u[:, 0] * (np.log(hi) - np.log(lo)))).astype(np.int32)
    rate = 0.1 + 4.9 * u[:, 1]
    decay = 1e-2 * u[:, 2]
    src = u[:, 3:].reshape(n_wells, sources_per_well, 3)
    sx = np.floor(src[:, :, 0] * side[:, None]).astype(np.int64)
    sy = np.floor(src[:, :, 1] * side[:, None]).astype(np.int64)
    well = np.repeat(np.arange(n_wells, dtype=np.int32), sources_per_well)
    cell = (sx * side[:, None] + sy).ravel()
    value = src[:, :, 2].ravel().astype(np.float32)
    return dict(xdims=side, ydims=side.copy(), rate=rate, decay=decay, src_well=well, src_cell=cell, src_val=value)


def mode_mixture(n, u, modes=8):
    """Sum of `modes` plane sine waves w*sin(2 pi (kx x + ky y)/n + phi) with
    kx, ky = 1 + floor(8U), w = 0.5 + 0.5U, phi = 2 pi U drawn in that order
    from u (5 values per mode), scaled to a maximum magnitude of 1."""
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    m = np.zeros((n, n))
    for k in range(modes):
        kx, ky, w, ph = 1 + np.floor(8 * u[5 * k]), 1 + np.floor(8 * u[5 * k + 1]), 0.5 + 0.5 * u[5 * k + 2], \
            2 * np.pi * u[5 * k + 3]
        m += w * np.sin(2 * np.pi * (kx * x + ky * y) / n + ph)
    return m / np.abs(m).max()


def adjoint_problem(n=1024, ticks=200, every=20, seed=42):
    """Adjoint workload: D = 1 + 0.5 M1 (M1 an 8-mode mixture from Rng(seed),
    so D in [0.5, 1.5]), decay 5e-3, 200 ticks with a checkpoint every 20;
    the target diffusivity is D * (1 + 0.05 M2) with a second mixture, and
    the initial field 16 Gaussian blobs (sigma 20) from Rng(seed + 1)."""
    u = halhip.rng_doubles(seed, 80)
    D = 1.0 + 0.5 * mode_mixture(n, u[:40])
    Dt = D * (1.0 + 0.05 * mode_mixture(n, u[40:]))
    u0 = adi_blobs(n=n, blobs=16, sigma=20.0, seed=seed + 1)["field"]
    return dict(X=n, Y=n, D=D.ravel(), D_target=Dt.ravel(), u0=u0, decay=5e-3, ticks=ticks, every=every)
