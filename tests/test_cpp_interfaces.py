"""The C++ drop-in boundary: Hal::Hip backends behind HAL's strategy
interfaces (include/Hal/Hip/HipBackends.h), exercised by model-side C++ code
(tests/cpp/hal_interface_tests.cpp)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
BIN = os.path.join(ROOT, "tests", "cpp", "hal_interface_tests")
LIB = os.path.join(ROOT, "hal-cpp_amd", "libhal_hip_backends.so")


def test_backends_library_exports_interface_classes():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for sym in ("Hal::Hip::HipDiffusionBackend2D::Diffuse", "Hal::Hip::HipAdiDiffusionBackend2D::DiffuseADI",
                "Hal::Hip::HipAgentForceBackend2D::AccumulateForces", "Hal::Internal::CheckFieldSizes",
                "Hal::Hip::HipContext::TryCreate", "Hal::Logger::Instance", "Hal::Hip::HipPDEGrid2D::Update",
                "Hal::Hip::HipPDEGrid2D::DiffusionADI"):
        assert sym in out, sym
    # vtables of the interface implementations are present
    assert re.search(r"vtable for Hal::Hip::HipAdiDiffusionBackend2D", out)


def test_cpp_program_skips_cleanly_without_gpu():
    import halhip
    if halhip.device_count() > 0:
        pytest.skip("host has a GPU")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 77, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_model_code_through_interfaces(tmp_path):
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
