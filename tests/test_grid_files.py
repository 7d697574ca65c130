"""HAL PDEGrid2D files (src/Grids/PDEGridSerialization.cpp:33-61): the
committed fixtures were written by the compiled reference's SavePDEGrid2D;
their layout is checked here byte by byte, and against the reference's own
loader when it is built."""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

HEADER = np.dtype([("x", "<i4"), ("y", "<i4"), ("wx", "u1"), ("wy", "u1"), ("ct", "<i4")])


def parse(path):
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:HEADER.itemsize], HEADER)[0]
    n = int(h["x"]) * int(h["y"])
    body = np.frombuffer(raw[HEADER.itemsize:], "<f8")
    assert body.size == 2 * n
    return int(h["x"]), int(h["y"]), int(h["wx"]), int(h["wy"]), int(h["ct"]), body[:n], body[n:]


def test_fixture_layout():
    X, Y, wx, wy, ct, field, deltas = parse(os.path.join(GOLDEN_DIR, "pdegrid2d_20x18.bin"))
    assert (X, Y, wx, wy, ct) == (20, 18, 1, 0, 37)
    assert HEADER.itemsize == 14
    np.testing.assert_array_equal(field.astype(np.float32).astype(np.float64), field)  # FP32-exact values
    assert np.all(deltas == 0.0)
    *_, f2, d2 = parse(os.path.join(GOLDEN_DIR, "pdegrid2d_20x18_pending.bin"))
    np.testing.assert_array_equal(f2, field)
    assert np.all(d2 == 0.5)


def test_reference_loader_reads_the_fixture():
    from oracle import Reference, available_reference
    if not available_reference():
        pytest.skip("compiled reference not built")
    got = Reference().load_pdegrid2d(os.path.join(GOLDEN_DIR, "pdegrid2d_20x18.bin"), 1000)
    X, Y, wx, wy, ct, field, deltas = parse(os.path.join(GOLDEN_DIR, "pdegrid2d_20x18.bin"))
    assert got[:5] == (X, Y, wx, wy, ct)
    np.testing.assert_array_equal(got[5], field)
    np.testing.assert_array_equal(got[6], deltas)
