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
