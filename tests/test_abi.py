"""CPU checks of the C ABI: the library loads, exports every function that
include/halhip.h declares, and fails cleanly without a GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

import halhip

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "halhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(halhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    halhip.load()
    out = subprocess.run(["nm", "-D", "--defined-only", halhip.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (halhip_[a-z0-9_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(halhip.exported_symbols()) <= set(declared_functions())


def test_rng_stream_matches_reference_golden(golden):
    np.testing.assert_array_equal(halhip.rng_doubles(42, 64), golden["rng42_doubles"])
    np.testing.assert_array_equal(halhip.rng_doubles(7, 16), golden["rng7_doubles"])


def test_rng_state_is_hal_rng_save_layout(golden):
    # Hal::Rng::Save of Rng(42) after 10 draws (src/Rng.cpp:122-125), written
    # by the compiled reference: the same 32 bytes, and restoring them
    # continues the reference's stream.
    r = halhip.Rng(42)
    np.testing.assert_array_equal(r.doubles(10), golden["rng42_doubles"][:10])
    assert r.save_bytes() == golden["rng42_state_after10"].tobytes()
    resumed = halhip.Rng(state=golden["rng42_state_after10"].tobytes())
    np.testing.assert_array_equal(resumed.doubles(54), golden["rng42_doubles"][10:])
    np.testing.assert_array_equal(r.doubles(54), golden["rng42_doubles"][10:])


def test_rng_state_round_trips_through_reference(tmp_path):
    from oracle import Reference, available_reference
    if not available_reference():
        pytest.skip("compiled reference not built")
    ref = Reference()
    # a state saved by HAL resumes here ...
    path = str(tmp_path / "hal_rng.bin")
    ref.rng_save(7, 5, path)
    with open(path, "rb") as fh:
        mine = halhip.Rng(state=fh.read())
    np.testing.assert_array_equal(mine.doubles(11), ref.rng_doubles(7, 16)[5:])
    # ... and a state saved here resumes in HAL
    r = halhip.Rng(99)
    r.doubles(3)
    with open(path, "wb") as fh:
        fh.write(r.save_bytes())
    np.testing.assert_array_equal(ref.rng_load_doubles(path, 8), r.doubles(8))


def test_null_arguments_are_invalid_argument():
    lib = halhip.load()
    assert lib.halhip_context_create(0, None) == halhip.ERR_INVALID_ARGUMENT
    assert "null" in lib.halhip_last_error().decode()
    assert lib.halhip_adi2d_upload(None, None, 0) == halhip.ERR_INVALID_ARGUMENT
    assert lib.halhip_diffusion2d_run_steps(None, 1, 0.1, 0.0, 0, 0, None) == halhip.ERR_INVALID_ARGUMENT


@pytest.mark.skipif(halhip.device_count() > 0, reason="host has a GPU")
def test_no_device_is_reported_not_crashed():
    with pytest.raises(halhip.NoDevice):
        halhip.Context(0)
