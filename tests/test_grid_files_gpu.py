"""Resident plans saved to / loaded from HAL PDEGrid2D files
(SavePDEGrid2D / LoadPDEGrid2D, src/Grids/PDEGridSerialization.cpp:33-61),
against fixtures written by the compiled reference."""
import os

import numpy as np
import pytest

import halhip
from conftest import GOLDEN_DIR
from test_grid_files import parse

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(GOLDEN_DIR, "pdegrid2d_20x18.bin")


@pytest.mark.parametrize("kind", ["explicit", "adi"])
def test_load_and_save_match_reference_files(ctx, tmp_path, kind):
    plan = (halhip.Diffusion2D if kind == "explicit" else halhip.Adi2D)(ctx, 20, 18)
    assert plan.load_grid(FIXTURE) == (True, False, 37)
    *_, field, _ = parse(FIXTURE)
    np.testing.assert_array_equal(plan.download().astype(np.float64), field)
    out = tmp_path / "saved.bin"
    plan.save_grid(out, True, False, 37)
    assert open(out, "rb").read() == open(FIXTURE, "rb").read()


def test_refusals(ctx, tmp_path):
    plan = halhip.Diffusion2D(ctx, 20, 18)
    with pytest.raises(halhip.InvalidArgument, match="pending deltas"):
        plan.load_grid(os.path.join(GOLDEN_DIR, "pdegrid2d_20x18_pending.bin"))
    other = halhip.Adi2D(ctx, 18, 20)
    with pytest.raises(halhip.InvalidArgument, match="saved grid is 20x18"):
        other.load_grid(FIXTURE)
    cut = tmp_path / "cut.bin"
    cut.write_bytes(open(FIXTURE, "rb").read()[:1000])
    with pytest.raises(halhip.InvalidArgument, match="failed read"):
        plan.load_grid(cut)
    with pytest.raises(halhip.InvalidArgument):
        plan.load_grid(tmp_path / "missing.bin")


def test_checkpoint_resume_is_exact(ctx, tmp_path):
    # run 30 ticks straight vs 12 ticks, save, load into a new plan, 18 ticks
    X, Y = 64, 48
    f = np.random.default_rng(3).random(X * Y).astype(np.float32)
    a = halhip.Adi2D(ctx, X, Y)
    a.upload(f)
    a.run_steps(30, 2.0, 0.0, False, False, bc=0.1)
    b = halhip.Adi2D(ctx, X, Y)
    b.upload(f)
    b.run_steps(12, 2.0, 0.0, False, False, bc=0.1)
    path = tmp_path / "ckpt.bin"
    b.save_grid(path, False, False, 12)
    c = halhip.Adi2D(ctx, X, Y)
    assert c.load_grid(path) == (False, False, 12)
    c.run_steps(18, 2.0, 0.0, False, False, bc=0.1)
    np.testing.assert_array_equal(c.download(), a.download())
