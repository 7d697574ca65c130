"""Batched wells on the GPU vs the FP64 oracle run well by well
(PDEGrid2D::DiffusionADI + MulAll + Add + Update per well), and vs the
per-instance launch loop of the same kernels."""
import numpy as np
import pytest

import halhip
import workloads
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def oracle_wells(oracle, batch, fields, w, ticks, wrap=(0, 0), bc=None):
    out = []
    for k in range(len(batch["xdims"])):
        X, Y = int(batch["xdims"][k]), int(batch["ydims"][k])
        sel = batch["src_well"] == k
        f = w.well(fields, k).astype(np.float64)
        out.append(oracle.adi_ticks(f, ticks, float(batch["rate"][k]), float(batch["decay"][k]), X, Y, *wrap, bc=bc,
                                    src_idx=batch["src_cell"][sel], src_val=batch["src_val"][sel].astype(np.float64)))
    return out


CARRIES = [halhip.ADI_CARRIES_SCAN, halhip.ADI_CARRIES_WINDOW]


def make(ctx, spec, rows, carries=halhip.ADI_CARRIES_SCAN):
    b = halhip.Batch(ctx, spec["xdims"], spec["ydims"])
    b.set_wells(spec["rate"], spec["decay"])
    b.set_sources(spec["src_well"], spec["src_cell"], spec["src_val"])
    b.set_strip_rows(rows)
    b.set_carries(carries)
    return b


@pytest.mark.parametrize("carries", CARRIES)
@pytest.mark.parametrize("rows", [4, 8])
def test_small_batch_vs_oracle(ctx, oracle, rows, carries):
    spec = workloads.wells_batch(n_wells=12, seed=7, lo=5, hi=300, sources_per_well=6)
    b = make(ctx, spec, rows, carries)
    rng = np.random.default_rng(1)
    f0 = rng.random(b.cells).astype(np.float32)
    b.upload(f0)
    b.run_ticks(12)
    got = b.download()
    ref = oracle_wells(oracle, spec, f0, b, 12)
    assert b.carries() == carries
    for k in range(len(spec["xdims"])):
        err = rel_l2(b.well(got, k), ref[k])
        assert err < 1e-5, (k, spec["xdims"][k], err)


@pytest.mark.parametrize("carries", CARRIES)
def test_wide_wells_and_boundaries(ctx, oracle, carries):
    # widths that select 1, 2 and 4 columns per thread, vector and scalar rows
    xd = np.array([3, 40, 17, 70, 9], np.int32)
    yd = np.array([256, 300, 512, 1000, 1024], np.int32)
    spec = dict(xdims=xd, ydims=yd, rate=np.array([0.3, 5.0, 2.0, 1.0, 4.0]),
                decay=np.array([0.0, 1e-3, 5e-3, 0.0, 1e-2]),
                src_well=np.array([0, 1, 1, 3, 4], np.int32),
                src_cell=np.array([5, 0, 299 + 300 * 39, 70 * 1000 - 1, 4000], np.int64),
                src_val=np.array([1.0, 0.5, 0.25, 2.0, 1.5], np.float32))
    b = make(ctx, spec, 8, carries)
    f0 = np.random.default_rng(2).random(b.cells).astype(np.float32)
    for wrap, bc in (((0, 0), None), ((1, 0), 0.2), ((0, 1), None)):
        b.upload(f0)
        b.run_ticks(7, *wrap, bc=bc)
        got = b.download()
        # a periodic x axis has no windowed carries: the batch keeps the scan
        want = halhip.ADI_CARRIES_SCAN if wrap[0] else carries
        assert b.carries() == want, (wrap, b.carries())
        ref = oracle_wells(oracle, spec, f0, b, 7, wrap, bc)
        for k in range(5):
            assert rel_l2(b.well(got, k), ref[k]) < 1e-5, (k, wrap, bc)


@pytest.mark.parametrize("rows", [8, 16, 4])
@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("carries", CARRIES)
def test_every_workgroup_shape_vs_oracle(ctx, oracle, carries, shape, rows):
    # every width class's workgroup shape option (threads x columns per
    # thread), with widths at the class edges and ragged vector groups
    xd = np.array([9, 20, 13, 33, 11, 7, 5], np.int32)
    yd = np.array([64, 255, 257, 510, 513, 1022, 1024], np.int32)
    spec = dict(xdims=xd, ydims=yd, rate=np.array([0.3, 5.0, 2.0, 1.0, 4.0, 0.7, 2.5]),
                decay=np.array([0.0, 1e-3, 5e-3, 0.0, 1e-2, 2e-3, 0.0]),
                src_well=np.array([0, 1, 2, 3, 4, 5, 6], np.int32),
                src_cell=np.array([5, 254, 257 * 12 + 3, 509, 600, 1021, 4 * 1024 + 1023], np.int64),
                src_val=np.array([1.0, 0.5, 0.25, 2.0, 1.5, 0.75, 1.25], np.float32))
    b = make(ctx, spec, 8, carries)
    b.set_shapes([shape] * 3)
    # 16-row strips need at most 2 columns per thread: the widest class's
    # default shape (4 columns) keeps 8
    cls_rows = [rows if (rows != 16 or cpt <= 2) else 8 for (_, _, cpt) in b.shapes()]
    b.set_class_rows(cls_rows)
    f0 = np.random.default_rng(12).random(b.cells).astype(np.float32)
    for wrap, bc in (((0, 0), None), ((0, 1), 0.3)):
        b.upload(f0)
        b.run_ticks(6, *wrap, bc=bc)
        got = b.download()
        ref = oracle_wells(oracle, spec, f0, b, 6, wrap, bc)
        for k in range(len(xd)):
            assert rel_l2(b.well(got, k), ref[k]) < 1e-5, (k, shape, cls_rows, wrap, bc)
    assert [s[0] for s in b.shapes()] == [shape] * 3
    assert b.class_rows() == cls_rows
    b.close()


def test_shape_and_rows_misuse(ctx, oracle):
    # out-of-range shape options and strip heights are refused and change
    # nothing; 16-row strips with a 4-columns-per-thread shape are refused
    # by the next run, before any tick
    xd = np.array([9, 20, 11], np.int32)
    yd = np.array([200, 400, 900], np.int32)
    spec = dict(xdims=xd, ydims=yd, rate=np.array([0.3, 2.0, 4.0]), decay=np.array([0.0, 1e-3, 5e-3]),
                src_well=np.array([0, 2], np.int32), src_cell=np.array([5, 900 * 5 + 7], np.int64),
                src_val=np.array([1.0, 0.5], np.float32))
    b = make(ctx, spec, 8)
    for bad in ([0, 3, 0], [-1, 0, 0], [0, 0]):
        with pytest.raises(halhip.InvalidArgument):
            b.set_shapes(bad)
    for bad in ([8, 12, 8], [8, 8, 0], [8, 8]):
        with pytest.raises(halhip.InvalidArgument):
            b.set_class_rows(bad)
    assert [s[0] for s in b.shapes()] == [0, 0, 0] and b.class_rows() == [8, 8, 8]
    f0 = np.random.default_rng(21).random(b.cells).astype(np.float32)
    b.upload(f0)
    b.set_strip_rows(16)   # the widest class's default shape: 256 threads x 4 columns
    with pytest.raises(halhip.InvalidArgument, match="16-row"):
        b.run_ticks(3)
    np.testing.assert_array_equal(b.download(), f0)
    b.set_shapes([0, 0, 1])   # 512 threads x 2 columns
    b.run_ticks(3)
    got = b.download()
    ref = oracle_wells(oracle, spec, f0, b, 3)
    for k in range(3):
        assert rel_l2(b.well(got, k), ref[k]) < 1e-5, k
    b.close()


@pytest.mark.parametrize("carries", CARRIES)
def test_per_instance_loop_matches_batched(ctx, carries):
    spec = workloads.wells_batch(n_wells=40, seed=3, lo=16, hi=512)
    b = make(ctx, spec, 8, carries)
    b.upload(np.zeros(b.cells, np.float32))
    b.run_ticks(25)
    batched = b.download()
    b.upload(np.zeros(b.cells, np.float32))
    b.run_ticks(25, per_instance=True)
    single = b.download()
    np.testing.assert_array_equal(batched, single)


def test_autotune_keeps_state_and_caches(ctx, tmp_path):
    spec = workloads.wells_batch(n_wells=20, seed=5, lo=32, hi=400)
    b = make(ctx, spec, 8)
    f0 = np.random.default_rng(3).random(b.cells).astype(np.float32)
    b.upload(f0)
    cache = str(tmp_path / "tune.txt")
    rows = b.autotune(5, cache_path=cache)
    assert rows in (4, 8, 16)
    np.testing.assert_array_equal(b.download(), f0)
    # "windowed:shape.rows,shape.rows,shape.rows"
    win, per = open(cache).read().split()[-1].split(":")
    pairs = [tuple(int(v) for v in p.split(".")) for p in per.split(",")]
    assert win in ("0", "1") and len(pairs) == 3
    assert [s[0] for s in b.shapes()] == [p[0] for p in pairs]
    assert b.class_rows() == [p[1] for p in pairs] and rows == pairs[2][1]
    assert b.autotune(5, cache_path=cache) == rows
    b.run_ticks(1)
    assert b.carries() == (halhip.ADI_CARRIES_WINDOW if win == "1" else halhip.ADI_CARRIES_SCAN)
    got = b.download()
    # the tuned configuration against the default one, tick for tick
    d = make(ctx, spec, 8)
    d.upload(f0)
    d.run_ticks(1)
    assert rel_l2(got, d.download()) < 1e-6


def test_workload_512_wells_first_ticks(ctx, oracle):
    spec = workloads.wells_batch()
    b = make(ctx, spec, 8)
    b.run_ticks(3)
    got = b.download()
    # check a sample of wells (small, median, largest) against the oracle
    order = np.argsort(spec["xdims"])
    worst = 0.0
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        X = int(spec["xdims"][k])
        sel = spec["src_well"] == k
        ref = oracle.adi_ticks(np.zeros(X * X), 3, float(spec["rate"][k]), float(spec["decay"][k]), X, X, 0, 0,
                               src_idx=spec["src_cell"][sel], src_val=spec["src_val"][sel].astype(np.float64))
        err = rel_l2(b.well(got, k), ref)
        worst = max(worst, err)
        assert err < 1e-5, (k, X)
    print(f"512 wells, 3 ticks (smallest, median, largest well): worst rel L2 vs FP64 = {worst:.3e}")
