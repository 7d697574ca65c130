"""halhip_agents2d_bins_noalloc: the bins of caller-owned device positions,
computed with no device allocation (a fixed per-GPU scratch array), equal
the resident plan's bins (upload + a forces-only binning + download_bins)
bit for bit, and the oracle's (CpuAgentForceBackend2D.cpp:85-109 order), on
two calls in a row, the second smaller, so the scratch array's leftovers from
the first call must not leak into the second."""
import ctypes

import numpy as np
import pytest

import halhip
import workloads

pytestmark = pytest.mark.gpu

_HIP_H2D, _HIP_D2H = 1, 2


def _hip():
    """The HIP runtime libhalhip.so itself is linked against (symbol lookup
    through a library handle searches its dependencies), so the test's buffers
    come from the same runtime instance as the library's work."""
    return halhip.load()


class DeviceBuffer:
    """A test-side device allocation (the entry point under test allocates nothing)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        rc = hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 4)))
        assert rc == 0, f"hipMalloc({nbytes}) -> {rc} {hip.hipGetErrorString(rc)}"
        self.nbytes = nbytes

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes),
                                  _HIP_H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes),
                                  _HIP_D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def plan_bins(ctx, pos, X, Y):
    n = pos.size // 2
    p = halhip.Agents2D(ctx, X, Y, n)
    p.upload(pos, np.full(n, 0.5, np.float32))
    p.run_ticks(0, 1.0, 1.0, 0.0, False, False)
    offsets, bucket = p.bins()
    p.close()
    return offsets, bucket


def test_bins_noalloc_matches_plan_and_oracle(ctx, oracle):
    hip = _hip()
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipGetErrorString.restype = ctypes.c_char_p
    hip.hipGetErrorString.argtypes = [ctypes.c_int]
    hip.hipGetLastError()
    # first call: 600k agents on 800 x 1000 bins (~0.75 per bin, crowded
    # bins included); second: 40k agents on 300 x 200 bins, denser
    for n, X, Y in ((600_000, 800, 1000), (40_000, 300, 200)):
        w = workloads.agents_uniform(n_agents=n, dim=1)
        pos = w["pos"].reshape(n, 2) * np.array([X, Y], np.float32)
        pos = pos.astype(np.float32).ravel()
        print(f"bins_noalloc: {n} agents on {X}x{Y} bins", flush=True)
        d_pos = DeviceBuffer(hip, pos.nbytes)
        d_off = DeviceBuffer(hip, 4 * (X * Y + 1))
        d_bkt = DeviceBuffer(hip, 4 * n)
        d_pos.put(pos)
        halhip.agents_bins_noalloc(ctx, d_pos.ptr.value, n, X, Y, d_off.ptr.value, d_bkt.ptr.value)
        ctx.synchronize()
        offsets = d_off.get(np.int32, X * Y + 1)
        bucket = d_bkt.get(np.int32, n)
        for b in (d_pos, d_off, d_bkt):
            b.free()
        want_off, want_bkt = plan_bins(ctx, pos, X, Y)
        assert np.array_equal(offsets, want_off)
        assert np.array_equal(bucket, want_bkt)
        ref_off, ref_bkt = oracle.agent_bins(pos.astype(np.float64), X, Y)
        assert np.array_equal(offsets, ref_off)
        assert np.array_equal(bucket, ref_bkt)


def test_bins_noalloc_rejects_sizes_beyond_its_scratch(ctx):
    with pytest.raises(halhip.InvalidArgument, match="MAX_BINS"):
        halhip.agents_bins_noalloc(ctx, 0, 0, 2048, 1024, 16, 16)
    with pytest.raises(halhip.InvalidArgument, match="MAX_AGENTS"):
        halhip.agents_bins_noalloc(ctx, 16, (1 << 20) + 1, 8, 8, 16, 16)
