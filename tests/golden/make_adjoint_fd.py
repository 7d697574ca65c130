"""Writes tests/golden/adjoint_fd_1024.npz: FP64 central finite differences of
the adjoint workload's loss (1024 x 1024, D = 1 + 0.5 * 8-mode mixture, decay
5e-3, 200 ticks; workloads.adjoint_problem) at 64 seeded cells and for the
decay rate.

The model is the FP64 oracle's per-cell-diffusivity ADI
(oracle_adi_cell_forward) evaluated at the FP32 values the device holds: D
and the decay rounded to FP32 and widened back, u0 as generated (FP32), and
the target u_T of the perturbed diffusivity D_target computed in FP64 and
rounded to FP32.  J = 1/2 |u_T - target|^2.  For each cell c,
    dJ/dD_c ~ (J(D + h e_c) - J(D - h e_c)) / (2 h),   h = 1e-6,
and likewise dJ/dk.  u_T is a degree-200 polynomial in k, so J''' is large
(~2e10 here) and the two-point difference in k carries a truncation error
h^2 J''' / 6 of ~1e-4 relative; the fourth-order difference
    (8 (J(k+h) - J(k-h)) - (J(k+2h) - J(k-2h))) / (12 h)
removes it and is stored as well (fd4_gradK).  Cells: x = floor(1024 U),
y = floor(1024 U) in order from Hal::Rng(7) (repeats skipped).  The loss of
the unperturbed model is stored too.  Run with the oracle built
(make -C oracle); 133 forward runs of 200 ticks, spread over worker
processes.
"""
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
for sub in ("oracle", "hal-cpp_amd"):
    sys.path.insert(0, os.path.join(ROOT, sub))

H = 1e-6
N_CELLS = 64


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def problem():
    import workloads
    from oracle import Oracle
    w = workloads.adjoint_problem()
    o = Oracle()
    o.set_threads(1)
    X, Y = w["X"], w["Y"]
    D = f32(w["D"])
    k = float(np.float32(w["decay"]))
    u0 = w["u0"].astype(np.float64)
    target = f32(o.adi_cell_forward(u0, f32(w["D_target"]), k, w["ticks"], X, Y))
    return w, X, Y, D, k, u0, target


def cells(X, Y):
    import halhip
    u = halhip.rng_doubles(7, 4 * N_CELLS)
    out = []
    for j in range(0, u.size, 2):
        c = int(np.floor(X * u[j])) * Y + int(np.floor(Y * u[j + 1]))
        if c not in out:
            out.append(c)
        if len(out) == N_CELLS:
            break
    assert len(out) == N_CELLS
    return np.array(out, np.int64)


_P = None


def _init():
    global _P
    _P = problem()


def _loss(job):
    """J for one perturbation: ("D", cell, sign) or ("k", -1, sign)."""
    from oracle import Oracle
    w, X, Y, D, k, u0, target = _P
    kind, c, sign = job
    Dv, kv = D, k
    if kind == "D":
        Dv = D.copy()
        Dv[c] += sign * H
    elif kind == "k":
        kv = k + sign * H
    o = Oracle()
    o.set_threads(1)
    uT = o.adi_cell_forward(u0, Dv, kv, w["ticks"], X, Y)
    return 0.5 * float(np.sum((uT - target) ** 2))


def main():
    w, X, Y, D, k, u0, target = problem()
    cs = cells(X, Y)
    jobs = [("J", -1, 0)]
    for c in cs:
        jobs += [("D", int(c), +1), ("D", int(c), -1)]
    jobs += [("k", -1, +1), ("k", -1, -1), ("k", -1, +2), ("k", -1, -2)]
    workers = max(1, min(16, os.cpu_count() or 1))
    with ProcessPoolExecutor(max_workers=workers, initializer=_init) as pool:
        losses = list(pool.map(_loss, jobs))
    J = losses[0]
    d = np.array(losses[1:1 + 2 * N_CELLS]).reshape(N_CELLS, 2)
    fd_D = (d[:, 0] - d[:, 1]) / (2 * H)
    kp1, km1, kp2, km2 = losses[-4:]
    fd_k = (kp1 - km1) / (2 * H)
    fd4_k = (8 * (kp1 - km1) - (kp2 - km2)) / (12 * H)
    out = os.path.join(HERE, "adjoint_fd_1024.npz")
    np.savez_compressed(out, cells=cs, fd_gradD=fd_D, fd_gradK=np.array([fd_k]), fd4_gradK=np.array([fd4_k]),
                        loss=np.array([J]),
                        h=np.array([H]), meta=np.array([X, Y, w["ticks"], w["every"]], np.int64))
    print(f"wrote {out}: J = {J:.9e}, dJ/dk = {fd_k:.6e} (fourth order {fd4_k:.6e}), |dJ/dD| at cells in "
          f"[{np.abs(fd_D).min():.3e}, {np.abs(fd_D).max():.3e}]")


if __name__ == "__main__":
    main()
