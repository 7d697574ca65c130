#!/usr/bin/env python3
"""Headline benchmark: implicit ADI diffusion on a 4096 x 4096 FP32 grid.

Workload: D = 5.0 (rate 5 at dt = dx = 1), Dirichlet boundary value 0, 64
seeded Gaussian blobs (sigma 20) as the initial field (workloads.adi_blobs).
One step = one ADI tick (x-implicit sweep then y-implicit sweep) over the
whole grid, inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 200] [--warmup 20]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hal-cpp_amd"))

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E
BYTES_PER_CELL_TICK = 16         # read u + write u* (x sweep), read u* + write u' (y sweep)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(field, n, rate, ticks):
    """The compiled HAL-CPP reference (PDEGrid2D::DiffusionADI + Update, one
    thread, -O2) timed on this host on `ticks` ticks of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Reference, available_reference
    import numpy as np
    if not available_reference():
        return None
    secs, _ = Reference().adi_ticks_timed(field.astype(np.float64), 1, ticks, rate, n, n, 0.0)
    return {"value": round(n * n * ticks / secs / 1e9, 6), "unit": "Gcell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{ticks} ticks (after 1 warm-up tick) of the same 4096^2 Dirichlet ADI workload, "
                      f"{secs:.2f} s; HAL-CPP PDEGrid2D::DiffusionADI compiled -O2, single-threaded "
                      f"(the reference CPU path has no threading); host {cpu_model()}, nproc {os.cpu_count()}"}


def traffic_from_profile():
    """HBM bytes per launch of the strip kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), as
    recorded under profiles/."""
    p = os.path.join(ROOT, "profiles", "adi_strip_tick_4096_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def reference_lib():
    """The compiled HAL-CPP reference (oracle/_ref), or None when not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Reference, available_reference
    return Reference() if available_reference() else None


def host_note():
    return f"host {cpu_model()}, nproc {os.cpu_count()}, one thread"


def bench_adi_model(ctx, w, ticks=200):
    """The model tick on the ADI workload's grid: DiffusionADI + MulAll(-1e-3)
    + Add at the 64 blob centres + Update, as HAL model code runs it, on the
    automatic path (decay and sources inside the LDS tick kernel) and on the
    register row solve (reduce, scan, tick and source launches)."""
    import halhip
    n = w["X"]
    u = halhip.rng_doubles(42, 3 * 64)
    idx = [int(n * u[3 * b]) * n + int(n * u[3 * b + 1]) for b in range(64)]
    out = {"workload": f"ADI {n}x{n}, D=5, Dirichlet 0, decay 1e-3, 64 point sources 0.01, {ticks} ticks"}
    for name, rows in (("fused", halhip.ADI_AUTO), ("register", halhip.ADI_ROWS_WORKGROUP_SCAN)):
        plan = halhip.Adi2D(ctx, n, n)
        plan.set_methods(rows, halhip.ADI_AUTO)
        plan.set_sources(idx, [0.01] * 64)
        plan.upload(w["field"])
        plan.run_steps(20, w["rate"], 1e-3, False, False, bc=w["bc"])
        plan.prepare_steps(ticks, w["rate"], 1e-3, False, False, bc=w["bc"])
        plan.sync()
        plan.run_steps(ticks, w["rate"], 1e-3, False, False, bc=w["bc"], asynchronous=True)
        ms = plan.sync()
        out[name + "_us_per_tick"] = round(ms * 1e3 / ticks, 3)
        out[name + "_Gcell_s"] = round(n * n * ticks / (ms * 1e-3) / 1e9, 2)
        plan.close()
    return out


def bench_long_rows(ctx, ticks=40, cpu=True, X=4096, Y=16384):
    """ADI on rows longer than one workgroup: the blob workload on a 4096 x
    16384 grid, Dirichlet 0, r = 5.  Rows are cut into 4096-column segments
    (reduce, cross-strip scan, segment solve, per-row segment scan and
    fix-up launches per tick).  Priced at the same 16 B/cell-tick as the
    whole-row tick."""
    import halhip
    import workloads
    w = workloads.adi_blobs(n=X, ny=Y)
    plan = halhip.Adi2D(ctx, X, Y)
    plan.upload(w["field"])
    plan.run_steps(5, w["rate"], 0.0, False, False, bc=w["bc"])
    plan.prepare_steps(ticks, w["rate"], 0.0, False, False, bc=w["bc"])
    plan.sync()
    plan.run_steps(ticks, w["rate"], 0.0, False, False, bc=w["bc"], asynchronous=True)
    ms = plan.sync()
    plan.set_profiling(True)
    plan.run_steps(10, w["rate"], 0.0, False, False, bc=w["bc"])
    prof = plan.profile()
    plan.set_profiling(False)
    cells = X * Y
    kernel_ms = sum(v[0] for v in prof.values()) / 10
    out = {"workload": f"ADI {X}x{Y} (rows of {Y} columns, 4096-column segments), D=5, Dirichlet 0, "
                       f"{ticks} ticks",
           "ms_per_tick": round(ms / ticks, 5), "Gcell_s": round(cells * ticks / (ms * 1e-3) / 1e9, 2),
           "kernels_us_per_launch": {k: round(v[0] / v[1] * 1e3, 2) for k, v in prof.items()},
           "kernel_ms_per_tick": round(kernel_ms, 5),
           "roofline": {"bound": "hbm", "achieved": round(BYTES_PER_CELL_TICK * cells / (kernel_ms * 1e-3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(BYTES_PER_CELL_TICK * cells / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "methods": list(plan.methods()), "strip_shape": list(plan.strip_shape())}
    plan.close()
    return out


def bench_explicit(ctx, ticks=1000, cpu=True, shape=None):
    """Explicit 256^2 (workloads.explicit_256): r = 0.2, decay 1e-3, centre
    source 1.0, periodic, 1000 ticks resident; 8 B/cell-tick algorithmic."""
    import halhip
    import workloads
    w = workloads.explicit_256()
    p = halhip.Diffusion2D(ctx, w["X"], w["Y"])
    if shape is not None:
        p.set_multi_shape(*shape)
    p.set_sources(w["src_idx"], w["src_val"])
    p.upload(w["field"])
    p.run_steps(100, w["rate"], w["decay"], True, True)
    p.upload(w["field"])
    p.prepare_steps(ticks, w["rate"], w["decay"], True, True)
    p.run_steps(ticks, w["rate"], w["decay"], True, True, asynchronous=True)
    ms = p.sync()
    cells = w["X"] * w["Y"]
    p.set_profiling(True)
    p.run_steps(100, w["rate"], w["decay"], True, True)
    prof = p.profile()
    multi = p.multi_shape()
    p.close()
    name, (total_ms, launches) = max(prof.items(), key=lambda kv: kv[1][0])
    out = {"workload": "explicit 256^2, r=0.2, decay 1e-3, centre source 1.0, periodic, 1000 ticks",
           "ms_total": round(ms, 4), "us_per_tick": round(ms * 1e3 / ticks, 3),
           "Gcell_s": round(cells * ticks / (ms * 1e-3) / 1e9, 2), "kernel": name,
           "kernel_us_per_launch": round(total_ms / launches * 1e3, 3),
           "ticks_per_launch": round(100 / launches, 2),
           "multi_shape": list(multi),
           "algorithmic_GBps": round(8 * cells * ticks / (ms * 1e-3) / 1e9, 1),
           "note": "256 KiB field stays in L2: several ticks per launch in LDS (temporal blocking); "
                   "8 B/cell-tick priced against HBM is not a bound here"}
    ref = reference_lib() if cpu else None
    if ref is not None:
        import numpy as np
        secs, _ = ref.grid_ticks_timed(w["field"].astype(np.float64), 1, ticks, False, w["rate"], w["decay"],
                                       w["X"], w["Y"], True, True, None, w["src_idx"],
                                       w["src_val"].astype(np.float64))
        out["cpu_baseline"] = {"value": round(cells * ticks / secs / 1e9, 5), "unit": "Gcell-updates/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"the full 1000 ticks, {secs:.2f} s: PDEGrid2D Diffusion + MulAll + Add + "
                                         f"Update compiled -O2; {host_note()}"}
        out["speedup_vs_cpu"] = round(out["Gcell_s"] / out["cpu_baseline"]["value"], 1)
    return out


def bench_agents(ctx, ticks=100, cpu=True):
    """4M agents on 2048^2 (workloads.agents_uniform): bin, scan, stable
    scatter, forces, move per tick; 52 N + 20 C bytes algorithmic per tick."""
    import halhip
    import workloads
    w = workloads.agents_uniform()
    n = w["rad"].size
    cells = w["X"] * w["Y"]
    p = halhip.Agents2D(ctx, w["X"], w["Y"], n)
    p.upload(w["pos"], w["rad"])
    p.run_ticks(10, w["R"], w["k"], w["dt"], False, False)
    p.upload(w["pos"], w["rad"])
    p.prepare_ticks(ticks, w["R"], w["k"], w["dt"], False, False)
    p.run_ticks(ticks, w["R"], w["k"], w["dt"], False, False, asynchronous=True)
    ms = p.sync()
    p.set_profiling(True)
    p.run_ticks(10, w["R"], w["k"], w["dt"], False, False)
    prof = p.profile()
    p.close()
    per_tick = ms / ticks
    alg = 52 * n + 20 * cells
    out = {"workload": "4M agents, 2048^2 bins, r=0.5, R=1, k=1, dt=0.1, no wrap, 100 ticks",
           "ms_per_tick": round(per_tick, 4), "Magent_ticks_s": round(n / (per_tick * 1e-3) / 1e6, 1),
           "algorithmic_bytes_per_tick": alg, "hbm_frac": round(alg / (per_tick * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
           "kernels_us_per_launch": {k: round(v[0] / v[1] * 1e3, 2) for k, v in prof.items()}}
    ref = reference_lib() if cpu else None
    if ref is not None:
        import numpy as np
        secs = ref.agent_forces_timed(w["pos"].astype(np.float64), w["rad"].astype(np.float64), 1, w["R"], w["k"],
                                      w["X"], w["Y"], False, False)
        out["cpu_baseline"] = {"value": round(n / secs / 1e6, 3), "unit": "Magent-ticks/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"one AccumulateForces call on the same 4M agents after one untimed call, "
                                         f"{secs:.2f} s (binning + forces; the move is not timed); {host_note()}"}
        out["speedup_vs_cpu"] = round(out["Magent_ticks_s"] / out["cpu_baseline"]["value"], 1)
    return out


def bench_batched(ctx, ticks=500, cpu=True):
    """512 uneven wells (workloads.wells_batch): batched ticks vs the same
    kernels launched well by well."""
    import numpy as np
    import halhip
    import workloads
    spec = workloads.wells_batch()
    b = halhip.Batch(ctx, spec["xdims"], spec["ydims"])
    b.set_wells(spec["rate"], spec["decay"])
    b.set_sources(spec["src_well"], spec["src_cell"], spec["src_val"])
    zeros = np.zeros(b.cells, np.float32)
    b.upload(zeros)
    b.autotune(20)  # strip rows x carry method x each width class's workgroup shape, timed on this batch
    b.run_ticks(5)
    b.prepare_ticks(ticks)
    b.run_ticks(ticks, asynchronous=True)
    batched_ms = b.sync() / ticks
    b.upload(zeros)
    b.run_ticks(2, per_instance=True)
    loop_ticks = 10
    b.run_ticks(loop_ticks, per_instance=True, asynchronous=True)
    loop_ms = b.sync() / loop_ticks
    plans = single_grid_wells(ctx, spec)
    out = {"workload": "512 wells, sides log-uniform 64-1024, D~U[0.1,5], decay~U[0,1e-2], 32 sources/well, "
                       "zero flux", "cells": int(b.cells), "strip_rows": b.strip_rows(),
           "carries": "windowed" if b.carries() == halhip.ADI_CARRIES_WINDOW else "scan",
           "workgroup_shapes": {f"width<={w}": f"{t} threads x {c} columns, {r}-row strips"
                                for w, (_, t, c), r in zip((256, 512, 1024), b.shapes(), b.class_rows())},
           "batched_ms_per_tick": round(batched_ms, 4), "per_instance_ms_per_tick": round(loop_ms, 4),
           "batched_Gcell_s": round(b.cells / (batched_ms * 1e-3) / 1e9, 2),
           "per_instance_Gcell_s": round(b.cells / (loop_ms * 1e-3) / 1e9, 2),
           "speedup": round(loop_ms / batched_ms, 2)}
    out.update(plans)
    out["speedup_vs_single_grid_plans"] = round(min(plans["plans_sequential_ms_per_tick"],
                                                    plans["plans_concurrent_ms_per_tick"]) / batched_ms, 2)
    # 16 B/cell-tick algorithmic, as for the single grid
    achieved = 16 * b.cells / (batched_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4)}
    b.close()
    ref = reference_lib() if cpu else None
    if ref is not None:
        import numpy as np
        secs = 0.0
        for wi in range(len(spec["xdims"])):
            side = int(spec["xdims"][wi])
            m = spec["src_well"] == wi
            s, _ = ref.grid_ticks_timed(np.zeros(side * side), 0, 1, True, spec["rate"][wi], spec["decay"][wi], side,
                                        side, False, False, None, spec["src_cell"][m],
                                        spec["src_val"][m].astype(np.float64))
            secs += s
        out["cpu_baseline"] = {"value": round(out["cells"] / secs / 1e9, 5), "unit": "Gcell-updates/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"one tick of every well in turn (MultiWellExperiment::Step shape), "
                                         f"{secs:.2f} s: PDEGrid2D DiffusionADI + MulAll + Add + Update; "
                                         f"{host_note()}"}
        out["speedup_vs_cpu"] = round(out["batched_Gcell_s"] / out["cpu_baseline"]["value"], 1)
        out["cpu_baseline_threads"] = cpu_wells_threaded(ref, spec, out["cells"])
    return out


def single_grid_wells(ctx, spec, ticks=20):
    """The stronger per-instance baseline: one single-grid ADI plan per well
    (the LDS row solve and windowed carries where the well's width allows),
    each replaying its own prepared 20-tick graph, run one well after
    another (sequential) and all enqueued at once on the plans' own streams
    (concurrent).  Same tick as the batch: ADI + decay + sources."""
    import numpy as np
    import halhip
    plans = []
    for w in range(len(spec["xdims"])):
        X, Y = int(spec["xdims"][w]), int(spec["ydims"][w])
        p = halhip.Adi2D(ctx, X, Y)
        m = spec["src_well"] == w
        p.set_sources(spec["src_cell"][m], spec["src_val"][m])
        p.upload(np.zeros(X * Y, np.float32))
        args = (float(spec["rate"][w]), float(spec["decay"][w]), False, False)
        p.run_steps(2, *args)
        p.prepare_steps(ticks, *args)
        plans.append((p, args))
    ctx.synchronize()
    t0 = time.perf_counter()
    for p, args in plans:
        p.run_steps(ticks, *args)  # returns when the well's ticks are done
    seq = (time.perf_counter() - t0) * 1e3 / ticks
    ctx.synchronize()
    t0 = time.perf_counter()
    for p, args in plans:
        p.run_steps(ticks, *args, asynchronous=True)
    ctx.synchronize()
    conc = (time.perf_counter() - t0) * 1e3 / ticks
    for p, _ in plans:
        p.close()
    return {"plans_sequential_ms_per_tick": round(seq, 4), "plans_concurrent_ms_per_tick": round(conc, 4)}


def cpu_wells_threaded(ref, spec, cells):
    """The reference's own parallel shape for a batch of wells:
    MultiWellExperiment::StepMultiThread runs one well per task on
    hardware_concurrency() threads, pulling wells from a shared counter
    (include/Hal/Tools/MultiWellExperiment.h:95-103,
    include/Hal/Concurrency/MultiThread.h:19-77).  The compiled reference's
    tick runs in a pool of the host's threads (ctypes releases the GIL); a
    step takes the busiest thread's summed tick time (grid set-up, which
    happens once per well in a model, is not counted)."""
    import collections
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    threads = max(1, min(16, os.cpu_count() or 1))
    busy = collections.defaultdict(float)
    lock = threading.Lock()

    def one(wi):
        side = int(spec["xdims"][wi])
        m = spec["src_well"] == wi
        s, _ = ref.grid_ticks_timed(np.zeros(side * side), 0, 1, True, spec["rate"][wi], spec["decay"][wi], side,
                                    side, False, False, None, spec["src_cell"][m],
                                    spec["src_val"][m].astype(np.float64))
        with lock:
            busy[threading.get_ident()] += s

    with ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(one, range(len(spec["xdims"]))))
    step = max(busy.values())
    return {"value": round(cells / step / 1e9, 5), "unit": "Gcell-updates/s", "cores": threads, "kind": "reference",
            "sample": f"one tick of every well, wells pulled by {threads} threads (StepMultiThread shape); "
                      f"busiest thread {step:.3f} s; host {cpu_model()}, nproc {os.cpu_count()}"}


def bench_per_call(ctx, cpu=True):
    """The per-call interfaces on host FP64 arrays, the reference's drop-in
    shape: IDiffusionBackend2D::Diffuse and IAdiDiffusionBackend2D::DiffuseADI
    (include/Hal/Pde/IDiffusionBackend2D.h:26-32,
    include/Hal/Pde/IAdiDiffusionBackend2D.h:25-31).  Every call converts to
    FP32, uploads, runs one step, downloads and accumulates into deltas; it is
    timed through the ctypes binding.  The CPU side is the reference's
    CpuDiffusionBackend2D / CpuAdiDiffusionBackend2D on the same arrays."""
    import numpy as np
    import halhip
    import workloads
    ref = reference_lib() if cpu else None
    cases = []
    for n in (256, 1024):
        f = np.random.default_rng(n).random(n * n)
        for adi in (False, True):
            plan = halhip.Adi2D(ctx, n, n) if adi else halhip.Diffusion2D(ctx, n, n)
            rate = 0.8 if adi else 0.2
            d = np.zeros(n * n)
            calls = 100 if n <= 256 else 20
            # as many untimed calls first: the first calls of a process pay
            # one-time host costs
            for _ in range(calls):
                plan.diffuse(f, d, rate, True, True)
            t0 = time.perf_counter()
            for _ in range(calls):
                plan.diffuse(f, d, rate, True, True)
            gpu_ms = (time.perf_counter() - t0) / calls * 1e3
            plan.close()
            case = {"interface": "DiffuseADI" if adi else "Diffuse", "grid": n, "wrap": "periodic",
                    "gpu_ms_per_call": round(gpu_ms, 4)}
            if ref is not None:
                cpu_calls = 20 if n <= 256 else 4
                cpu_ms = ref.diffuse_calls_timed(f, rate, n, n, True, True, adi, cpu_calls) / cpu_calls * 1e3
                case["cpu_ms_per_call"] = round(cpu_ms, 4)
                case["speedup_vs_cpu"] = round(cpu_ms / gpu_ms, 2)
            cases.append(case)
    # IAgentForceBackend2D::AccumulateForces (include/Hal/Agents/IAgentForceBackend2D.h:28-36):
    # upload, bin, forces and the download of the velocity deltas in every call
    aw = workloads.agents_uniform(n_agents=1_000_000, dim=1024)
    pos = aw["pos"].astype(np.float64)
    rad = aw["rad"].astype(np.float64)
    n_ag = rad.size
    plan = halhip.Agents2D(ctx, aw["X"], aw["Y"], n_ag)
    for _ in range(3):
        plan.accumulate_forces(pos, rad, aw["R"], aw["k"], 0, 0)
    t0 = time.perf_counter()
    for _ in range(10):
        plan.accumulate_forces(pos, rad, aw["R"], aw["k"], 0, 0)
    gpu_ms = (time.perf_counter() - t0) / 10 * 1e3
    plan.close()
    case = {"interface": "AccumulateForces", "agents": n_ag, "bins": [aw["X"], aw["Y"]], "gpu_ms_per_call": round(gpu_ms, 4)}
    if ref is not None:
        cpu_ms = ref.agent_forces_timed(pos, rad, 1, aw["R"], aw["k"], aw["X"], aw["Y"], 0, 0) * 1e3
        case["cpu_ms_per_call"] = round(cpu_ms, 4)
        case["speedup_vs_cpu"] = round(cpu_ms / gpu_ms, 2)
    cases.append(case)
    return {"workload": "per-call interfaces on FP64 host arrays: Diffuse / DiffuseADI periodic, r = 0.2 (explicit) "
                        "and 0.8 (ADI); AccumulateForces on 1M agents over 1024^2 bins (r = 0.5, R = 1, k = 1)",
            "cases": cases,
            "note": "each call includes the FP64->FP32 conversion, upload, download and the FP64 result "
                    "(accumulated into deltas for the diffusion interfaces, written for the forces); "
                    "CPU: the reference backends on one thread"}


def bench_adjoint(ctx, cpu=True):
    """1024^2 per-cell-diffusivity ADI (workloads.adjoint_problem): forward
    200 ticks vs forward + checkpointed adjoint gradient."""
    import halhip
    import workloads
    w = workloads.adjoint_problem()
    p = halhip.Adjoint2D(ctx, w["X"], w["Y"])
    p.set_problem(w["D_target"], w["decay"], w["ticks"], w["every"])
    target = p.forward(w["u0"])
    p.set_problem(w["D"], w["decay"], w["ticks"], w["every"])
    p.forward(w["u0"])
    t0 = time.perf_counter()
    p.forward(w["u0"])
    fwd = time.perf_counter() - t0
    p.gradient(w["u0"], target)
    t0 = time.perf_counter()
    p.gradient(w["u0"], target)
    grad = time.perf_counter() - t0
    p.close()
    # The plain model on the same grid: uniform-rate ADI ticks (rate 1, the
    # mean of D) with the same decay, resident, through the HAL-interface path.
    plain = halhip.Adi2D(ctx, w["X"], w["Y"])
    plain.upload(w["u0"])
    plain.run_steps(20, 1.0, w["decay"], False, False)
    plain.sync()
    plain.run_steps(w["ticks"], 1.0, w["decay"], False, False, asynchronous=True)
    plain_ms = plain.sync()
    plain.close()
    cells = w["X"] * w["Y"]
    fwd_gbps = 24 * cells * w["ticks"] / fwd / 1e9
    grad_gbps = 104 * cells * w["ticks"] / grad / 1e9
    out = {"workload": "1024^2, D = 1 + 0.5 * 8-mode mixture, decay 5e-3, 200 ticks, checkpoint every 20",
           "forward_ms": round(fwd * 1e3, 3), "gradient_ms": round(grad * 1e3, 3),
           "gradient_over_forward": round(grad / fwd, 3),
           "forward_Gcell_s": round(cells * w["ticks"] / fwd / 1e9, 2),
           "plain_adi_ms": round(plain_ms, 3),
           "gradient_over_plain": round(grad * 1e3 / plain_ms, 2),
           "note": "forward and gradient wall times include the 4 MiB upload of u0 (and readback of u_T); "
                   "plain_adi_ms is the uniform-rate resident ADI on the same grid",
           "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                        "forward_achieved": round(fwd_gbps, 1), "forward_frac": round(fwd_gbps / HBM_PEAK_GBS, 4),
                        "gradient_achieved": round(grad_gbps, 1),
                        "gradient_frac": round(grad_gbps / HBM_PEAK_GBS, 4),
                        "bytes": "24 B/cell-tick forward, 104 B/cell-tick gradient"}}
    if cpu:
        # The reference has no adjoint: the FP64 restatement's adjoint is
        # the CPU figure (10 ticks, scaled to the 200-tick problem).
        import numpy as np
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        sample = 10
        t0 = time.perf_counter()
        Oracle().adi_cell_gradient(w["u0"].astype(np.float64), w["D"], w["decay"], sample, w["X"], w["Y"],
                                   target.astype(np.float64))
        secs = (time.perf_counter() - t0) * w["ticks"] / sample
        out["cpu_baseline"] = {"value": round(secs * 1e3, 1), "unit": "ms per gradient", "cores": 1, "kind": "port",
                               "sample": f"FP64 CPU adjoint (forward with all intermediates stored + adjoint sweep) "
                                         f"over {sample} ticks, scaled x{w['ticks'] // sample}; {host_note()}"}
        out["speedup_vs_cpu"] = round(secs / grad, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-ticks", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adi-only", action="store_true", help="skip the batched and adjoint comparisons")
    ap.add_argument("--size", type=int, default=4096,
                    help="side of the ADI workload grid (default 4096, the headline configuration)")
    ap.add_argument("--shape", default=None,
                    help="strip shape columns_per_thread,strip_rows,vector(0|1) instead of the automatic one")
    ap.add_argument("--methods", default=None,
                    help="ADI row solver and carries as ROWS,CARRIES with ROWS in auto|scan|lds and "
                         "CARRIES in auto|scan|window, instead of the automatic choice")
    ap.add_argument("--wrap", default="0,0",
                    help="periodic axes as WRAPX,WRAPY (0|1); a non-periodic axis keeps the Dirichlet 0 edges")
    ap.add_argument("--explicit-shape", default=None,
                    help="rows per workgroup and ticks per launch of the explicit companion as B,T")
    ap.add_argument("--companion", choices=["explicit", "agents", "batched", "adjoint", "per_call", "long_rows"],
                    default=None,
                    help="run only this companion workload and print its JSON (for profiling)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run the copy kernel (read + write this many MiB, 10 launches) for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import halhip
    import workloads

    if args.calibrate_copy_mib:
        ctx = halhip.Context(0)
        nbytes = args.calibrate_copy_mib << 20
        ms = ctx.calibration_copy(nbytes, 10)
        print(json.dumps({"calibration_copy_bytes": nbytes, "ms_per_copy": ms,
                          "GBps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return

    if args.companion:
        ctx = halhip.Context(0)
        run = {"explicit": bench_explicit, "agents": bench_agents, "batched": bench_batched,
               "adjoint": bench_adjoint, "per_call": bench_per_call, "long_rows": bench_long_rows}[args.companion]
        kwargs = {}
        if args.companion == "explicit" and args.explicit_shape:
            kwargs["shape"] = tuple(int(v) for v in args.explicit_shape.split(","))
        print(json.dumps({args.companion: run(ctx, cpu=not args.no_cpu_baseline, **kwargs)}), flush=True)
        ctx.close()
        return

    wx, wy = (int(v) != 0 for v in args.wrap.split(","))
    headline = not wx and not wy
    w = workloads.adi_blobs(n=args.size)
    n = w["X"]
    cells = n * n
    ctx = halhip.Context(0)
    plan = halhip.Adi2D(ctx, n, n)
    if args.shape:
        cpt, rows, vec = (int(v) for v in args.shape.split(","))
        plan.set_strip_shape(cpt, rows, bool(vec))
    if args.methods:
        rows_name, carries_name = args.methods.split(",")
        plan.set_methods({"auto": halhip.ADI_AUTO, "scan": halhip.ADI_ROWS_WORKGROUP_SCAN,
                          "lds": halhip.ADI_ROWS_LDS}[rows_name],
                         {"auto": halhip.ADI_AUTO, "scan": halhip.ADI_CARRIES_SCAN,
                          "window": halhip.ADI_CARRIES_WINDOW}[carries_name])
    plan.upload(w["field"])

    plan.run_steps(args.warmup, w["rate"], 0.0, wx, wy, bc=w["bc"])
    # Capture and instantiate, before the clock starts, every graph the timed
    # run replays (whatever --warmup left the reduction half at).
    plan.prepare_steps(args.steps, w["rate"], 0.0, wx, wy, bc=w["bc"])
    ctx.synchronize()
    t0 = time.perf_counter()
    plan.run_steps(args.steps, w["rate"], 0.0, wx, wy, bc=w["bc"], asynchronous=True)
    event_ms = plan.sync()
    ctx.synchronize()
    wall = time.perf_counter() - t0

    r_used, c_used, k_used = plan.methods()
    adi_methods = ({halhip.ADI_ROWS_WORKGROUP_SCAN: "scan", halhip.ADI_ROWS_LDS: "lds"}.get(r_used, r_used),
                   {halhip.ADI_CARRIES_SCAN: "scan", halhip.ADI_CARRIES_WINDOW: "window"}.get(c_used, c_used), k_used)

    # Per-kernel device time (HIP events around each launch on the plan's
    # stream), in a separate profiled pass of the same ticks.
    plan.set_profiling(True)
    prof_ticks = min(args.steps, 50)
    plan.run_steps(prof_ticks, w["rate"], 0.0, wx, wy, bc=w["bc"])
    prof = plan.profile()
    plan.set_profiling(False)
    dom = "adi_strip_tick"
    dom_ms = prof[dom][0] / prof[dom][1]
    achieved = BYTES_PER_CELL_TICK * cells / (dom_ms * 1e-3) / 1e9
    traffic = traffic_from_profile() if n == 4096 and headline else None

    extras = {}
    if not args.shape and not args.methods and headline:
        extras["adi_model_tick"] = bench_adi_model(ctx, w)
    if not args.adi_only:
        cpu = not args.no_cpu_baseline
        extras["explicit"] = bench_explicit(ctx, cpu=cpu)
        extras["agents"] = bench_agents(ctx, cpu=cpu)
        extras["batched"] = bench_batched(ctx, cpu=cpu)
        extras["adjoint"] = bench_adjoint(ctx, cpu=cpu)
        extras["per_call"] = bench_per_call(ctx, cpu=cpu)
        extras["long_rows"] = bench_long_rows(ctx, cpu=cpu)

    base = None
    if not args.no_cpu_baseline and headline:
        base = cpu_baseline(w["field"], n, w["rate"], args.cpu_ticks)

    value = cells * args.steps / wall / 1e9
    edges = {(False, False): "Dirichlet 0", (True, True): "periodic x and y",
             (True, False): "periodic x, Dirichlet 0 in y", (False, True): "Dirichlet 0 in x, periodic y"}[(wx, wy)]
    # The headline metric names four figures: the ADI rate (value), its
    # share of the HBM peak, and the batched and adjoint paths against the
    # plain one; the last three are also top-level keys below.
    full = headline and n == 4096 and not args.adi_only
    metric = ("Gcell-updates/s and % HBM peak (ADI 4096\u00b2); batched & adjoint vs plain" if full
              else f"Gcell-updates/s (ADI {n}^2 FP32, {edges})")
    line = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 64 Gaussian blobs (sigma 20, amplitude U[0.5,1]) at positions from Hal::Rng(42)",
        "config": {"workload": f"ADI {n}x{n}, D=5, dt=1, {edges}, {args.steps} ticks", "grid": [n, n],
                   "bc": edges, "rate": w["rate"], "parallelism": "single GPU",
                   "strip_shape": list(plan.strip_shape()),
                   "methods": dict(zip(("rows", "carries", "window"), adi_methods))},
        "hbm_frac_wall": round(value * BYTES_PER_CELL_TICK / HBM_PEAK_GBS, 4),
        "event_ms_per_step": round(event_ms / args.steps, 5),
        "kernels_ms_per_launch": {k: round(v[0] / v[1], 5) for k, v in prof.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom, "algorithmic_bytes_per_launch": BYTES_PER_CELL_TICK * cells},
        "cpu_baseline": base,
    }
    line.update(extras)
    line["hbm_peak_pct"] = round(100.0 * achieved / HBM_PEAK_GBS, 2)
    if "batched" in extras:
        line["batched_vs_per_instance"] = extras["batched"]["speedup"]
    if "adjoint" in extras:
        line["adjoint_gradient_vs_plain"] = extras["adjoint"]["gradient_over_plain"]
    if base is not None:
        line["speedup_vs_cpu"] = round(value / base["value"], 1)
    print(json.dumps(line), flush=True)
    plan.close()
    ctx.close()


if __name__ == "__main__":
    main()
