"""Writes tests/golden/*.npz from the compiled HAL-CPP reference.

Run in a container that has the reference sources (it builds
oracle/_ref/libhalref.so through `make -C oracle ref`).  The fixtures hold
inputs and the reference's outputs only; tests compare the FP64 oracle and the
HIP backend against them.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True)
    from oracle import Reference
    ref = Reference()
    out = {}

    # Hal::Rng streams (src/Rng.cpp).
    out["rng42_doubles"] = ref.rng_doubles(42, 64)
    out["rng1234_u64"] = ref.rng_u64(1234, 16)
    out["rng7_doubles"] = ref.rng_doubles(7, 16)
    # Hal::Rng::Save of Rng(42) after 10 draws (src/Rng.cpp:122-125): 32 raw bytes
    state_path = os.path.join(HERE, "_rng_state.tmp")
    ref.rng_save(42, 10, state_path)
    with open(state_path, "rb") as fh:
        out["rng42_state_after10"] = np.frombuffer(fh.read(), np.uint8).copy()
    os.remove(state_path)

    rng = np.random.default_rng(20240601)
    cases = []
    # (X, Y, wrapX, wrapY, bc-kind) : bc-kind 0 none, 1 constant, 2 edge arrays
    shapes = [(1, 1), (1, 5), (4, 1), (2, 3), (12, 18), (20, 14), (17, 33)]
    k = 0
    for (X, Y) in shapes:
        for wx in (0, 1):
            for wy in (0, 1):
                for bck in (0, 1, 2):
                    if bck and wx and wy:
                        continue
                    f = rng.random(X * Y)
                    d = rng.random(X * Y) * 0.1
                    if bck == 0:
                        bc = None
                    elif bck == 1:
                        bc = float(rng.random())
                    else:
                        bc = (rng.random(Y), rng.random(Y), rng.random(X), rng.random(X))
                    for op, rate in (("diffuse", 0.2), ("adi", 0.6), ("adi", 5.0)):
                        if op == "diffuse":
                            res = ref.diffuse(f, d, rate, X, Y, wx, wy, bc)
                        else:
                            res = ref.adi(f, d, rate, X, Y, wx, wy, bc)
                        pre = f"case{k}_"
                        out[pre + "meta"] = np.array([X, Y, wx, wy, bck, 0 if op == "diffuse" else 1], np.int64)
                        out[pre + "rate"] = np.array([rate])
                        out[pre + "field"] = f
                        out[pre + "deltas_in"] = d
                        if bck == 1:
                            out[pre + "bc"] = np.array([bc])
                        elif bck == 2:
                            out[pre + "bc_xlo"], out[pre + "bc_xhi"], out[pre + "bc_ylo"], out[pre + "bc_yhi"] = bc
                        out[pre + "deltas_out"] = res
                        k += 1
    out["ncases"] = np.array([k])

    # Per-cell-rate explicit diffusion (src/Pde/PDEequations.cpp:199-228).
    X, Y = 9, 11
    f = rng.random(X * Y)
    rx, ry = 0.05 + 0.2 * rng.random(X * Y), 0.05 + 0.2 * rng.random(X * Y)
    for wx in (0, 1):
        for wy in (0, 1):
            out[f"rates_{wx}{wy}_out"] = ref.diffuse_rates(f, np.zeros(X * Y), rx, ry, X, Y, wx, wy)
    out["rates_field"], out["rates_x"], out["rates_y"] = f, rx, ry

    # Full PDEGrid2D tick compositions.
    f = rng.random(16 * 16)
    out["ticks_explicit_field"] = f
    out["ticks_explicit_out"] = ref.grid_ticks(f, 25, 0, 0.2, 1e-3, 16, 16, 1, 1, None, [8 * 16 + 8], [1.0])
    f = rng.random(24 * 20)
    out["ticks_adi_field"] = f
    out["ticks_adi_out"] = ref.grid_ticks(f, 6, 1, 5.0, 5e-3, 24, 20, 0, 0, 0.25, [3, 77, 77, 400], [1.0, 0.5, 0.25, 2.0])
    out["ticks_adi_zf_out"] = ref.grid_ticks(f, 6, 1, 0.8, 0.0, 24, 20, 0, 1, None, [], [])

    # Agent forces (src/Agents/CpuAgentForceBackend2D.cpp), scene like
    # tests/IAgentForceBackend2DTests.cpp:25-37 from Hal::Rng(1234).
    u = ref.rng_doubles(1234, 180)
    pos = np.empty(120)
    rad = np.empty(60)
    for i in range(60):
        pos[2 * i] = u[3 * i] * 16.0
        pos[2 * i + 1] = u[3 * i + 1] * 16.0
        rad[i] = 0.4 + u[3 * i + 2] * 0.4
    out["agents_pos"], out["agents_rad"] = pos, rad
    out["agents_wrap_out"] = ref.agent_forces(pos, rad, 3.0, 0.4, 16, 16, 1, 1)
    out["agents_nowrap_out"] = ref.agent_forces(pos, rad, 3.0, 0.4, 16, 16, 0, 0)
    pos2 = rng.random(2 * 500) * np.tile([40.0, 24.0], 500)
    pos2[:20] = np.floor(pos2[:20])          # exact integers: floor boundaries
    rad2 = np.full(500, 0.5)
    out["agents2_pos"] = pos2
    out["agents2_out"] = ref.agent_forces(pos2, rad2, 1.0, 1.0, 40, 24, 0, 1)

    # Upwind advection, uniform and per-cell velocity
    # (src/Pde/PDEequations.cpp:380-413, 535-573; include/Hal/Pde/PDEequations.h:413-447, 721-783),
    # and model ticks Diffusion + Advection + MulAll + Add + Update.
    adv = 0
    for (X, Y) in [(1, 1), (1, 6), (5, 1), (7, 9), (16, 12)]:
        for wx in (0, 1):
            for wy in (0, 1):
                for bck in (0, 2):
                    f = rng.random(X * Y)
                    d = rng.random(X * Y) * 0.1
                    vel = rng.uniform(-0.25, 0.25, 2)
                    xv = rng.uniform(-0.25, 0.25, X * Y)
                    yv = rng.uniform(-0.25, 0.25, X * Y)
                    if bck:
                        bc = (rng.random(Y), rng.random(Y), rng.random(X), rng.random(X))
                        vbc = (rng.uniform(-0.25, 0.25, Y), rng.uniform(-0.25, 0.25, Y),
                               rng.uniform(-0.25, 0.25, X), rng.uniform(-0.25, 0.25, X))
                    else:
                        bc = vbc = None
                    pre = f"adv{adv}_"
                    out[pre + "meta"] = np.array([X, Y, wx, wy, bck], np.int64)
                    out[pre + "field"], out[pre + "deltas_in"], out[pre + "vel"] = f, d, vel
                    out[pre + "xv"], out[pre + "yv"] = xv, yv
                    if bck:
                        out[pre + "bc_xlo"], out[pre + "bc_xhi"], out[pre + "bc_ylo"], out[pre + "bc_yhi"] = bc
                        out[pre + "vbc_xlo"], out[pre + "vbc_xhi"], out[pre + "vbc_ylo"], out[pre + "vbc_yhi"] = vbc
                    out[pre + "uniform_out"] = ref.advect(f, d, vel[0], vel[1], X, Y, wx, wy, bc)
                    out[pre + "cells_out"] = ref.advect_cells(f, d, xv, yv, X, Y, wx, wy, bc, vbc)
                    adv += 1
    out["nadv"] = np.array([adv])
    f = rng.random(20 * 18)
    out["ticks_model_field"] = f
    out["ticks_model_out"] = ref.model_ticks(f, 15, 0.15, 0.2, -0.1, 2e-3, 20, 18, 1, 0, None, [5, 200], [1.0, 0.5])
    out["ticks_model_bc_out"] = ref.model_ticks(f, 15, 0.15, -0.05, 0.25, 0.0, 20, 18, 0, 0, 0.4, [7], [0.3])

    np.savez_compressed(os.path.join(HERE, "hal_reference_vectors.npz"), **out)

    # PDEGrid2D files written by Hal::SavePDEGrid2D (src/Grids/PDEGridSerialization.cpp:33-43):
    # FP32-exact field values (a resident FP32 plan saves them bit for bit),
    # zero deltas; and the same grid with pending deltas.
    f = rng.random(20 * 18).astype(np.float32).astype(np.float64)
    ref.save_pdegrid2d(os.path.join(HERE, "pdegrid2d_20x18.bin"), f, np.zeros(f.size), 20, 18, 1, 0, 37)
    ref.save_pdegrid2d(os.path.join(HERE, "pdegrid2d_20x18_pending.bin"), f, np.full(f.size, 0.5), 20, 18, 1, 0, 37)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
