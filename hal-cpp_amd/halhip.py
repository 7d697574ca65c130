"""ctypes binding of libhalhip.so, the C ABI declared in include/halhip.h.

This is the Python mirror of HAL-CPP's three strategy interfaces used by the
tests and the benchmark (the C++ mirror is include/Hal/Hip/*.h).  It loads the
in-tree library and raises when it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhalhip.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_NO_DEVICE = 2
ERR_OUT_OF_MEMORY = 3
ERR_HIP = 4
ERR_UNSUPPORTED = 5
BC_NONE = 0
BC_DIRICHLET = 1
ADI_MAX_Y = 8192
DELTA_SET = 0
DELTA_ADD = 1
DELTA_MUL = 2
DELTA_SCALE = 3
DELTA_SET_NON_NEGATIVE = 4
REDUCE_MAX_DELTA = 0
REDUCE_MAX_DELTA_SCALED = 1
REDUCE_FIELD_MAX = 2
REDUCE_FIELD_MIN = 3
REDUCE_FIELD_SUM = 4
ADI_AUTO = 0
ADI_ROWS_WORKGROUP_SCAN = 1
ADI_ROWS_LDS = 2
ADI_ROWS_SEGMENTED = 3
ADI_CARRIES_SCAN = 1
ADI_CARRIES_WINDOW = 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


class HalBC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("value", ctypes.c_double),
                ("xlo", _dp), ("xhi", _dp), ("ylo", _dp), ("yhi", _dp)]


class HalHipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class InvalidArgument(HalHipError, ValueError):
    pass


class NoDevice(HalHipError):
    pass


# Exported symbols, with argtypes.  Everything returns int.
_SIGS = {
    "halhip_last_error": None,
    "halhip_device_count": [_ip],
    "halhip_context_create": [_i32, ctypes.POINTER(_vp)],
    "halhip_context_destroy": [_vp],
    "halhip_context_fail_next_allocs": [_vp, _i32],
    "halhip_context_synchronize": [_vp],
    "halhip_calibration_copy": [_vp, _i64, _i32, _fp],
    "halhip_rng_doubles": [ctypes.c_uint64, _i64, _dp],
    "halhip_rng_seed_state": [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)],
    "halhip_rng_next_doubles": [ctypes.POINTER(ctypes.c_uint64), _i64, _dp],
    "halhip_diffusion2d_create": [_vp, _i32, _i32, ctypes.POINTER(_vp)],
    "halhip_diffusion2d_destroy": [_vp],
    "halhip_diffusion2d_diffuse_f64": [_vp, _dp, _dp, _i64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_diffuse_rates_f64": [_vp, _dp, _dp, _dp, _dp, _i64, _i32, _i32],
    "halhip_diffusion2d_advect_f64": [_vp, _dp, _dp, _i64, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_advect_cells_f64": [_vp, _dp, _dp, _dp, _dp, _i64, _i32, _i32, ctypes.POINTER(HalBC),
                                            ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_set_advection": [_vp, _f64, _f64],
    "halhip_diffusion2d_save_grid": [_vp, ctypes.c_char_p, _i32, _i32, _i32],
    "halhip_diffusion2d_load_grid": [_vp, ctypes.c_char_p, _ip, _ip, _ip],
    "halhip_adi2d_save_grid": [_vp, ctypes.c_char_p, _i32, _i32, _i32],
    "halhip_adi2d_load_grid": [_vp, ctypes.c_char_p, _ip, _ip, _ip],
    "halhip_diffusion2d_set_multi_shape": [_vp, _i32, _i32],
    "halhip_diffusion2d_get_multi_shape": [_vp, _ip, _ip],
    "halhip_diffusion2d_upload": [_vp, _fp, _i64],
    "halhip_diffusion2d_download": [_vp, _fp, _i64],
    "halhip_diffusion2d_set_sources": [_vp, _i32, _lp, _fp],
    "halhip_diffusion2d_run_steps": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_run_steps_async": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_sync": [_vp, _fp],
    "halhip_diffusion2d_prepare_steps": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_delta_point_ops": [_vp, _i32, _ip, _lp, _dp],
    "halhip_diffusion2d_delta_all": [_vp, _i32, _f64],
    "halhip_diffusion2d_delta_set_values": [_vp, _dp, _i64],
    "halhip_diffusion2d_delta_diffuse": [_vp, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_delta_diffuse_rates": [_vp, _dp, _dp, _i64, _i32, _i32],
    "halhip_diffusion2d_delta_advect": [_vp, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_delta_advect_cells": [_vp, _dp, _dp, _i64, _i32, _i32, ctypes.POINTER(HalBC),
                                              ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_delta_adi": [_vp, _vp, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_diffusion2d_delta_reduce": [_vp, _i32, _f64, _dp],
    "halhip_diffusion2d_delta_update": [_vp],
    "halhip_diffusion2d_download_deltas": [_vp, _fp, _i64],
    "halhip_diffusion2d_set_profiling": [_vp, _i32],
    "halhip_diffusion2d_profile": [_vp, _i32, _ip, ctypes.c_char_p, _i32, _dp, _lp],
    "halhip_adi2d_create": [_vp, _i32, _i32, ctypes.POINTER(_vp)],
    "halhip_adi2d_destroy": [_vp],
    "halhip_adi2d_diffuse_f64": [_vp, _dp, _dp, _i64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_adi2d_upload": [_vp, _fp, _i64],
    "halhip_adi2d_download": [_vp, _fp, _i64],
    "halhip_adi2d_set_sources": [_vp, _i32, _lp, _fp],
    "halhip_adi2d_run_steps": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_adi2d_run_steps_async": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_adi2d_sync": [_vp, _fp],
    "halhip_adi2d_prepare_steps": [_vp, _i32, _f64, _f64, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_adi2d_set_strip_shape": [_vp, _i32, _i32, _i32],
    "halhip_adi2d_get_strip_shape": [_vp, _ip, _ip, _ip],
    "halhip_adi2d_set_methods": [_vp, _i32, _i32],
    "halhip_adi2d_get_methods": [_vp, _ip, _ip, _ip],
    "halhip_adi2d_set_profiling": [_vp, _i32],
    "halhip_adi2d_profile": [_vp, _i32, _ip, ctypes.c_char_p, _i32, _dp, _lp],
    "halhip_batch_create": [_vp, _i32, _ip, _ip, ctypes.POINTER(_vp)],
    "halhip_batch_destroy": [_vp],
    "halhip_batch_set_wells": [_vp, _dp, _dp],
    "halhip_batch_set_sources": [_vp, _i32, _ip, _lp, _fp],
    "halhip_batch_upload": [_vp, _fp, _i64],
    "halhip_batch_download": [_vp, _fp, _i64],
    "halhip_batch_run_ticks": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_batch_run_ticks_async": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_batch_run_ticks_per_instance": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_batch_run_ticks_per_instance_async": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_batch_sync": [_vp, _fp],
    "halhip_batch_prepare_ticks": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_batch_set_strip_rows": [_vp, _i32],
    "halhip_batch_get_strip_rows": [_vp, _ip],
    "halhip_batch_set_carries": [_vp, _i32],
    "halhip_batch_get_carries": [_vp, _ip],
    "halhip_batch_autotune": [_vp, _i32, _i32, _i32, ctypes.POINTER(HalBC), ctypes.c_char_p, _ip],
    "halhip_batch_set_profiling": [_vp, _i32],
    "halhip_batch_profile": [_vp, _i32, _ip, ctypes.c_char_p, _i32, _dp, _lp],
    "halhip_batch_set_class_rows": [_vp, _ip],
    "halhip_batch_get_class_rows": [_vp, _ip],
    "halhip_agents2d_set_scan_spin_limit": [_vp, _i32],
    "halhip_batch_set_shapes": [_vp, _ip],
    "halhip_batch_get_shapes": [_vp, _ip, _ip, _ip],
    "halhip_adjoint2d_create": [_vp, _i32, _i32, ctypes.POINTER(_vp)],
    "halhip_adjoint2d_destroy": [_vp],
    "halhip_adjoint2d_set_problem": [_vp, _dp, _f64, _i32, _i32],
    "halhip_adjoint2d_set_problem_bc": [_vp, _dp, _f64, _i32, _i32, _i32, _i32, ctypes.POINTER(HalBC)],
    "halhip_adjoint2d_forward": [_vp, _fp, _fp],
    "halhip_adjoint2d_gradient": [_vp, _fp, _fp, _dp, _fp, _dp],
    "halhip_adjoint2d_set_profiling": [_vp, _i32],
    "halhip_adjoint2d_profile": [_vp, _i32, _ip, ctypes.c_char_p, _i32, _dp, _lp],
    "halhip_agents2d_create": [_vp, _i32, _i32, _i32, ctypes.POINTER(_vp)],
    "halhip_agents2d_destroy": [_vp],
    "halhip_agents2d_forces_f64": [_vp, _dp, _dp, _i32, _dp, _f64, _f64, _i32, _i32],
    "halhip_agents2d_upload": [_vp, _fp, _fp, _i32],
    "halhip_agents2d_download_positions": [_vp, _fp, _i32],
    "halhip_agents2d_download_velocities": [_vp, _fp, _i32],
    "halhip_agents2d_download_bins": [_vp, _ip, _ip],
    "halhip_agents2d_set_motion": [_vp, _f64, _f64],
    "halhip_agents2d_upload_motion": [_vp, _fp, _i32],
    "halhip_agents2d_download_motion": [_vp, _fp, _i32],
    "halhip_agents2d_run_ticks": [_vp, _i32, _f64, _f64, _f64, _i32, _i32],
    "halhip_agents2d_run_ticks_async": [_vp, _i32, _f64, _f64, _f64, _i32, _i32],
    "halhip_agents2d_sync": [_vp, _fp],
    "halhip_agents2d_prepare_ticks": [_vp, _i32, _f64, _f64, _f64, _i32, _i32],
    "halhip_agents2d_bins_noalloc": [_vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp],
    "halhip_agents2d_set_profiling": [_vp, _i32],
    "halhip_agents2d_profile": [_vp, _i32, _ip, ctypes.c_char_p, _i32, _dp, _lp],
}


def exported_symbols():
    return list(_SIGS)


_lib = None


def load(path=LIB_PATH):
    """Load libhalhip.so (once).  Raises FileNotFoundError when not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() or make -C hal-cpp_amd")
    lib = ctypes.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        if name == "halhip_last_error":
            fn.restype = ctypes.c_char_p
            fn.argtypes = []
            continue
        fn.restype = ctypes.c_int
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc == OK:
        return
    msg = load().halhip_last_error().decode()
    if rc == ERR_INVALID_ARGUMENT:
        raise InvalidArgument(rc, msg)
    if rc == ERR_NO_DEVICE:
        raise NoDevice(rc, msg)
    raise HalHipError(rc, msg)


def device_count():
    n = ctypes.c_int32(0)
    _check(load().halhip_device_count(ctypes.byref(n)))
    return n.value


def agents_bins_noalloc(ctx, positions_dev, n_agents, X, Y, offsets_dev, bucket_dev, stream=None):
    """halhip_agents2d_bins_noalloc on caller-owned DEVICE buffers (integer
    addresses): enqueues the binning on `stream` (None: the null stream) and
    returns without waiting."""
    _check(load().halhip_agents2d_bins_noalloc(ctx.h, _vp(positions_dev), n_agents, X, Y, _vp(offsets_dev),
                                               _vp(bucket_dev), _vp(stream or 0)))


def rng_doubles(seed, n):
    out = np.empty(n, np.float64)
    _check(load().halhip_rng_doubles(seed, n, out.ctypes.data_as(_dp)))
    return out


class Rng:
    """Hal::Rng as an explicit state: `state` (4 uint64) is byte for byte
    what Hal::Rng::Save writes (src/Rng.cpp:122-130)."""

    def __init__(self, seed=None, state=None):
        self.state = np.zeros(4, np.uint64)
        if state is not None:
            st = np.frombuffer(bytes(state), np.uint64) if isinstance(state, (bytes, bytearray)) else state
            self.state[:] = np.asarray(st, np.uint64)
        else:
            _check(load().halhip_rng_seed_state(seed, self.state.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))

    def doubles(self, n):
        out = np.empty(n, np.float64)
        _check(load().halhip_rng_next_doubles(self.state.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
                                              out.ctypes.data_as(_dp)))
        return out

    def save_bytes(self):
        """The 32 bytes of Hal::Rng::Save."""
        return self.state.tobytes()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def make_bc(bc, X, Y):
    """None, a constant Dirichlet value, or a 4-tuple of edge arrays
    (bc(-1,y), bc(X,y), bc(x,-1), bc(x,Y)).  Returns (struct-or-None, keepalive)."""
    if bc is None:
        return None, ()
    s = HalBC()
    s.kind = BC_DIRICHLET
    if np.isscalar(bc):
        s.value = float(bc)
        return s, ()
    arrs = tuple(_f64(a) for a in bc)
    _require(len(arrs) == 4 and arrs[0].size == Y and arrs[1].size == Y and arrs[2].size == X and arrs[3].size == X,
             "bc edge arrays must hold yDim, yDim, xDim and xDim values")
    s.xlo, s.xhi, s.ylo, s.yhi = (a.ctypes.data_as(_dp) for a in arrs)
    return s, arrs


def _check_field_sizes(field, deltas, X, Y, caller):
    """Internal::CheckFieldSizes (src/Pde/IDiffusionBackend2D.cpp:20-33): both
    arrays must hold xDim*yDim values; deltas is updated in place."""
    if not (isinstance(deltas, np.ndarray) and deltas.dtype == np.float64 and deltas.flags.c_contiguous):
        raise InvalidArgument(ERR_INVALID_ARGUMENT, caller + ": deltas must be a contiguous float64 array")
    if field.size != X * Y or deltas.size != X * Y:
        raise InvalidArgument(ERR_INVALID_ARGUMENT, caller + ": field and deltas must each hold xDim*yDim values")


def _require(ok, msg):
    if not ok:
        raise InvalidArgument(ERR_INVALID_ARGUMENT, msg)


def _bcp(s):
    return ctypes.byref(s) if s is not None else None


class Context:
    """halhip_context: replaces Hal::Gpu::ComputeContext."""

    def __init__(self, device=0):
        self.h = _vp()
        _check(load().halhip_context_create(device, ctypes.byref(self.h)))

    def fail_next_allocs(self, count):
        _check(load().halhip_context_fail_next_allocs(self.h, count))

    def calibration_copy(self, nbytes, repeats):
        ms = ctypes.c_float(0.0)
        _check(load().halhip_calibration_copy(self.h, nbytes, repeats, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(load().halhip_context_synchronize(self.h))

    def close(self):
        if self.h:
            load().halhip_context_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Plan:
    _prefix = ""

    def _fn(self, name):
        return getattr(load(), f"halhip_{self._prefix}_{name}")

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on):
        _check(self._fn("set_profiling")(self.h, 1 if on else 0))

    def profile(self):
        """{kernel_name: (total_ms, launches)} accumulated since profiling was enabled."""
        n = ctypes.c_int32(0)
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_int64 * 32)()
        _check(self._fn("profile")(self.h, 32, ctypes.byref(n), names, 4096, ms, cnt))
        keys = names.value.decode().split(",") if n.value else []
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def save_grid(self, path, wrapX, wrapY, update_ct=0):
        """Resident field to a HAL PDEGrid2D file (zero deltas)."""
        _check(self._fn("save_grid")(self.h, str(path).encode(), int(wrapX), int(wrapY), update_ct))

    def load_grid(self, path):
        """Resident field from a HAL PDEGrid2D file; returns (wrapX, wrapY, updateCt)."""
        wx, wy, ct = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._fn("load_grid")(self.h, str(path).encode(), ctypes.byref(wx), ctypes.byref(wy), ctypes.byref(ct)))
        return bool(wx.value), bool(wy.value), ct.value

    def sync(self):
        t = ctypes.c_float(0.0)
        _check(self._fn("sync")(self.h, ctypes.byref(t)))
        return t.value


class Diffusion2D(_Plan):
    """Explicit diffusion plan (IDiffusionBackend2D + resident RunSteps)."""
    _prefix = "diffusion2d"

    def __init__(self, ctx, X, Y):
        self.ctx, self.X, self.Y = ctx, X, Y
        self.h = _vp()
        _check(load().halhip_diffusion2d_create(ctx.h, X, Y, ctypes.byref(self.h)))

    def diffuse(self, field, deltas, rate, wrapX, wrapY, bc=None):
        """deltas += one explicit step (in place on the float64 `deltas`)."""
        field = _f64(field)
        _check_field_sizes(field, deltas, self.X, self.Y, type(self).__name__ + ".diffuse")
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("diffuse_f64")(self.h, field.ctypes.data_as(_dp), deltas.ctypes.data_as(_dp), field.size,
                                       rate, int(wrapX), int(wrapY), _bcp(s)))
        return deltas

    def diffuse_rates(self, field, deltas, ratesX, ratesY, wrapX, wrapY):
        field, ratesX, ratesY = _f64(field), _f64(ratesX), _f64(ratesY)
        _check_field_sizes(field, deltas, self.X, self.Y, "Diffusion2D.diffuse_rates")
        if ratesX.size != field.size or ratesY.size != field.size:
            raise InvalidArgument(ERR_INVALID_ARGUMENT, "Diffusion2D.diffuse_rates: rates must hold xDim*yDim values")
        _check(self._fn("diffuse_rates_f64")(self.h, field.ctypes.data_as(_dp), deltas.ctypes.data_as(_dp),
                                             ratesX.ctypes.data_as(_dp), ratesY.ctypes.data_as(_dp),
                                             field.size, int(wrapX), int(wrapY)))
        return deltas

    def advect(self, field, deltas, xVel, yVel, wrapX, wrapY, bc=None):
        """deltas += uniform-velocity upwind advection (in place on the float64 `deltas`)."""
        field = _f64(field)
        _check_field_sizes(field, deltas, self.X, self.Y, "Diffusion2D.advect")
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("advect_f64")(self.h, field.ctypes.data_as(_dp), deltas.ctypes.data_as(_dp), field.size,
                                      xVel, yVel, int(wrapX), int(wrapY), _bcp(s)))
        return deltas

    def advect_cells(self, field, deltas, xVels, yVels, wrapX, wrapY, bc=None, vel_bc=None):
        """deltas += per-cell-velocity upwind advection (flux form)."""
        field, xVels, yVels = _f64(field), _f64(xVels), _f64(yVels)
        _check_field_sizes(field, deltas, self.X, self.Y, "Diffusion2D.advect_cells")
        if xVels.size != field.size or yVels.size != field.size:
            raise InvalidArgument(ERR_INVALID_ARGUMENT, "Diffusion2D.advect_cells: velocities must hold xDim*yDim values")
        s, keep = make_bc(bc, self.X, self.Y)
        v, keepv = make_bc(vel_bc, self.X, self.Y)
        _check(self._fn("advect_cells_f64")(self.h, field.ctypes.data_as(_dp), deltas.ctypes.data_as(_dp),
                                            xVels.ctypes.data_as(_dp), yVels.ctypes.data_as(_dp), field.size,
                                            int(wrapX), int(wrapY), _bcp(s), _bcp(v)))
        return deltas

    def set_multi_shape(self, rows_per_group=0, ticks_per_launch=0):
        """Pin the multi-tick shape (0, 0: automatic)."""
        _check(self._fn("set_multi_shape")(self.h, rows_per_group, ticks_per_launch))

    def multi_shape(self):
        b, t = ctypes.c_int32(), ctypes.c_int32()
        _check(self._fn("get_multi_shape")(self.h, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    def set_advection(self, xVel, yVel):
        _check(self._fn("set_advection")(self.h, xVel, yVel))

    def upload(self, field):
        field = _f32(field)
        _check(self._fn("upload")(self.h, field.ctypes.data_as(_fp), field.size))

    def download(self):
        out = np.empty(self.X * self.Y, np.float32)
        _check(self._fn("download")(self.h, out.ctypes.data_as(_fp), out.size))
        return out

    def set_sources(self, idx, val):
        idx = np.ascontiguousarray(idx, np.int64)
        val = _f32(val)
        _check(self._fn("set_sources")(self.h, idx.size, idx.ctypes.data_as(_lp), val.ctypes.data_as(_fp)))

    def run_steps(self, steps, rate, decay, wrapX, wrapY, bc=None, asynchronous=False):
        s, keep = make_bc(bc, self.X, self.Y)
        name = "run_steps_async" if asynchronous else "run_steps"
        _check(self._fn(name)(self.h, steps, rate, decay, int(wrapX), int(wrapY), _bcp(s)))

    def prepare_steps(self, steps, rate, decay, wrapX, wrapY, bc=None):
        """Build the graphs run_steps(steps, ...) replays; runs no tick."""
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("prepare_steps")(self.h, steps, rate, decay, int(wrapX), int(wrapY), _bcp(s)))


    # PDEGrid2D delta calls on the resident field (include/Hal/Grids/PDEGrid2D.h:89-399)
    def delta_point_ops(self, ops, index, values):
        ops = np.ascontiguousarray(ops, np.int32)
        index = np.ascontiguousarray(index, np.int64)
        values = _f64(values)
        _check(self._fn("delta_point_ops")(self.h, ops.size, ops.ctypes.data_as(_ip), index.ctypes.data_as(_lp),
                                           values.ctypes.data_as(_dp)))

    def delta_all(self, op, value=0.0):
        _check(self._fn("delta_all")(self.h, op, value))

    def delta_set_values(self, values):
        values = _f64(values)
        _check(self._fn("delta_set_values")(self.h, values.ctypes.data_as(_dp), values.size))

    def delta_diffuse(self, rate, wrapX, wrapY, bc=None):
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("delta_diffuse")(self.h, rate, int(wrapX), int(wrapY), _bcp(s)))

    def delta_diffuse_rates(self, ratesX, ratesY, wrapX, wrapY):
        rx, ry = _f64(ratesX), _f64(ratesY)
        _check(self._fn("delta_diffuse_rates")(self.h, rx.ctypes.data_as(_dp), ry.ctypes.data_as(_dp), rx.size,
                                               int(wrapX), int(wrapY)))

    def delta_advect(self, xVel, yVel, wrapX, wrapY, bc=None):
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("delta_advect")(self.h, xVel, yVel, int(wrapX), int(wrapY), _bcp(s)))

    def delta_advect_cells(self, xVels, yVels, wrapX, wrapY, bc=None, vel_bc=None):
        xv, yv = _f64(xVels), _f64(yVels)
        s, keep = make_bc(bc, self.X, self.Y)
        v, keepv = make_bc(vel_bc, self.X, self.Y)
        _check(self._fn("delta_advect_cells")(self.h, xv.ctypes.data_as(_dp), yv.ctypes.data_as(_dp), xv.size,
                                              int(wrapX), int(wrapY), _bcp(s), _bcp(v)))

    def delta_adi(self, adi, rate, wrapX, wrapY, bc=None):
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("delta_adi")(self.h, adi.h, rate, int(wrapX), int(wrapY), _bcp(s)))

    def delta_reduce(self, what, offset=0.0):
        out = ctypes.c_double(0.0)
        _check(self._fn("delta_reduce")(self.h, what, offset, ctypes.byref(out)))
        return out.value

    def delta_update(self):
        _check(self._fn("delta_update")(self.h))

    def deltas(self):
        out = np.empty(self.X * self.Y, np.float32)
        _check(self._fn("download_deltas")(self.h, out.ctypes.data_as(_fp), out.size))
        return out


class Adi2D(_Plan):
    """ADI plan (IAdiDiffusionBackend2D + resident RunSteps)."""
    _prefix = "adi2d"

    def set_strip_shape(self, columns_per_thread, strip_rows, vector=True):
        _check(self._fn("set_strip_shape")(self.h, columns_per_thread, strip_rows, 1 if vector else 0))

    def strip_shape(self):
        c, s, v = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._fn("get_strip_shape")(self.h, ctypes.byref(c), ctypes.byref(s), ctypes.byref(v)))
        return c.value, s.value, bool(v.value)

    def set_methods(self, rows=ADI_AUTO, carries=ADI_AUTO):
        """Plain-tick methods (halhip_adi2d_set_methods); a method that does not
        apply falls back to the other."""
        _check(self._fn("set_methods")(self.h, rows, carries))

    def methods(self):
        """(row solver, carries, window K) used by the last plain run."""
        r, c, k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._fn("get_methods")(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(k)))
        return r.value, c.value, k.value

    def __init__(self, ctx, X, Y):
        self.ctx, self.X, self.Y = ctx, X, Y
        self.h = _vp()
        _check(load().halhip_adi2d_create(ctx.h, X, Y, ctypes.byref(self.h)))

    def diffuse(self, field, deltas, rate, wrapX, wrapY, bc=None):
        field = _f64(field)
        _check_field_sizes(field, deltas, self.X, self.Y, type(self).__name__ + ".diffuse")
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("diffuse_f64")(self.h, field.ctypes.data_as(_dp), deltas.ctypes.data_as(_dp), field.size,
                                       rate, int(wrapX), int(wrapY), _bcp(s)))
        return deltas

    def upload(self, field):
        field = _f32(field)
        _check(self._fn("upload")(self.h, field.ctypes.data_as(_fp), field.size))

    def download(self):
        out = np.empty(self.X * self.Y, np.float32)
        _check(self._fn("download")(self.h, out.ctypes.data_as(_fp), out.size))
        return out

    def set_sources(self, idx, val):
        idx = np.ascontiguousarray(idx, np.int64)
        val = _f32(val)
        _check(self._fn("set_sources")(self.h, idx.size, idx.ctypes.data_as(_lp), val.ctypes.data_as(_fp)))

    def run_steps(self, steps, rate, decay, wrapX, wrapY, bc=None, asynchronous=False):
        s, keep = make_bc(bc, self.X, self.Y)
        name = "run_steps_async" if asynchronous else "run_steps"
        _check(self._fn(name)(self.h, steps, rate, decay, int(wrapX), int(wrapY), _bcp(s)))

    def prepare_steps(self, steps, rate, decay, wrapX, wrapY, bc=None):
        """Build the graphs run_steps(steps, ...) replays; runs no tick."""
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("prepare_steps")(self.h, steps, rate, decay, int(wrapX), int(wrapY), _bcp(s)))


class Agents2D(_Plan):
    """Agent-force plan (IAgentForceBackend2D + resident ticks)."""
    _prefix = "agents2d"

    def __init__(self, ctx, X, Y, max_agents):
        self.ctx, self.X, self.Y, self.cap = ctx, X, Y, max_agents
        self.n = 0
        self.h = _vp()
        _check(load().halhip_agents2d_create(ctx.h, X, Y, max_agents, ctypes.byref(self.h)))

    def accumulate_forces(self, positions, radii, R, k, wrapX, wrapY):
        """IAgentForceBackend2D::AccumulateForces: returns the float64 (2N) deltas."""
        positions, radii = _f64(positions), _f64(radii)
        if positions.size != 2 * radii.size:
            raise InvalidArgument(ERR_INVALID_ARGUMENT, "positions must hold two values per radius")
        out = np.zeros(positions.size, np.float64)
        _check(self._fn("forces_f64")(self.h, positions.ctypes.data_as(_dp), radii.ctypes.data_as(_dp), radii.size,
                                      out.ctypes.data_as(_dp), R, k, int(wrapX), int(wrapY)))
        self.n = radii.size
        return out

    def upload(self, positions, radii):
        positions, radii = _f32(positions), _f32(radii)
        _require(positions.size == 2 * radii.size, "Agents2D.upload: positions must hold two values per radius")
        _check(self._fn("upload")(self.h, positions.ctypes.data_as(_fp), radii.ctypes.data_as(_fp), radii.size))
        self.n = radii.size

    def run_ticks(self, ticks, R, k, dt, wrapX, wrapY, asynchronous=False):
        name = "run_ticks_async" if asynchronous else "run_ticks"
        _check(self._fn(name)(self.h, ticks, R, k, dt, int(wrapX), int(wrapY)))

    def prepare_ticks(self, ticks, R, k, dt, wrapX, wrapY):
        """Build the graphs run_ticks(ticks, ...) replays; runs no tick."""
        _check(self._fn("prepare_ticks")(self.h, ticks, R, k, dt, int(wrapX), int(wrapY)))

    def positions(self):
        out = np.empty(2 * self.n, np.float32)
        _check(self._fn("download_positions")(self.h, out.ctypes.data_as(_fp), self.n))
        return out

    def velocities(self):
        out = np.empty(2 * self.n, np.float32)
        _check(self._fn("download_velocities")(self.h, out.ctypes.data_as(_fp), self.n))
        return out

    def set_scan_spin_limit(self, limit):
        """Polls a bin-scan tile makes for a predecessor's total before the
        scan gives up (reported as an error by the next download)."""
        _check(self._fn("set_scan_spin_limit")(self.h, limit))

    def set_motion(self, friction, max_vel=0.0):
        """Carry friction * velocity across ticks and cap its norm at max_vel
        (0: no cap), as SphericalAgent2D::ApplyFriction / CapVelocity."""
        _check(self._fn("set_motion")(self.h, friction, max_vel))

    def upload_motion(self, velocities):
        velocities = _f32(velocities)
        _require(velocities.size == 2 * self.n, "Agents2D.upload_motion: velocities must hold two values per agent")
        _check(self._fn("upload_motion")(self.h, velocities.ctypes.data_as(_fp), self.n))

    def motion(self):
        out = np.empty(2 * self.n, np.float32)
        _check(self._fn("download_motion")(self.h, out.ctypes.data_as(_fp), self.n))
        return out

    def bins(self):
        offsets = np.empty(self.X * self.Y + 1, np.int32)
        bucket = np.empty(max(self.n, 1), np.int32)
        _check(self._fn("download_bins")(self.h, offsets.ctypes.data_as(_ip), bucket.ctypes.data_as(_ip)))
        return offsets, bucket[:self.n]


class Batch(_Plan):
    """Batched wells of uneven size (one batched launch per phase per tick)."""
    _prefix = "batch"

    def __init__(self, ctx, xdims, ydims):
        xdims = np.ascontiguousarray(xdims, np.int32)
        ydims = np.ascontiguousarray(ydims, np.int32)
        self.ctx = ctx
        self.xdims, self.ydims = xdims, ydims
        self.cells = int((xdims.astype(np.int64) * ydims).sum())
        self.offsets = np.concatenate([[0], np.cumsum(xdims.astype(np.int64) * ydims)])
        self.h = _vp()
        _check(load().halhip_batch_create(ctx.h, xdims.size, xdims.ctypes.data_as(_ip), ydims.ctypes.data_as(_ip),
                                          ctypes.byref(self.h)))

    def set_wells(self, rates, decays=None):
        rates = _f64(rates)
        decays = None if decays is None else _f64(decays)
        _check(self._fn("set_wells")(self.h, rates.ctypes.data_as(_dp),
                                     None if decays is None else decays.ctypes.data_as(_dp)))

    def set_sources(self, well, cell, value):
        well = np.ascontiguousarray(well, np.int32)
        cell = np.ascontiguousarray(cell, np.int64)
        value = _f32(value)
        _check(self._fn("set_sources")(self.h, well.size, well.ctypes.data_as(_ip), cell.ctypes.data_as(_lp),
                                       value.ctypes.data_as(_fp)))

    def upload(self, fields):
        fields = _f32(fields)
        _check(self._fn("upload")(self.h, fields.ctypes.data_as(_fp), fields.size))

    def download(self):
        out = np.empty(self.cells, np.float32)
        _check(self._fn("download")(self.h, out.ctypes.data_as(_fp), out.size))
        return out

    def well(self, fields, w):
        return fields[self.offsets[w]:self.offsets[w + 1]]

    def run_ticks(self, ticks, wrapX=False, wrapY=False, bc=None, asynchronous=False, per_instance=False):
        s, keep = make_bc(bc, 1, 1) if (bc is None or np.isscalar(bc)) else (None, ())
        name = "run_ticks_per_instance" if per_instance else "run_ticks"
        if asynchronous:
            name += "_async"
        _check(self._fn(name)(self.h, ticks, int(wrapX), int(wrapY), _bcp(s)))

    def prepare_ticks(self, ticks, wrapX=False, wrapY=False, bc=None):
        """Build the graphs run_ticks(ticks, ...) replays; runs no tick."""
        s, keep = make_bc(bc, 1, 1) if (bc is None or np.isscalar(bc)) else (None, ())
        _check(self._fn("prepare_ticks")(self.h, ticks, int(wrapX), int(wrapY), _bcp(s)))

    def set_strip_rows(self, rows):
        _check(self._fn("set_strip_rows")(self.h, rows))

    def strip_rows(self):
        r = ctypes.c_int32()
        _check(self._fn("get_strip_rows")(self.h, ctypes.byref(r)))
        return r.value

    def set_carries(self, method):
        _check(self._fn("set_carries")(self.h, method))

    def set_shapes(self, shapes):
        """Workgroup shape option (0..2) of each width class (<= 256, <= 512,
        <= 1024 columns)."""
        a = np.ascontiguousarray(shapes, np.int32)
        _require(a.size == 3, "Batch.set_shapes: one option per width class (3)")
        _check(self._fn("set_shapes")(self.h, a.ctypes.data_as(_ip)))

    def set_class_rows(self, rows):
        """Strip height (4, 8 or 16) of each width class."""
        a = np.ascontiguousarray(rows, np.int32)
        _require(a.size == 3, "Batch.set_class_rows: one strip height per width class (3)")
        _check(self._fn("set_class_rows")(self.h, a.ctypes.data_as(_ip)))

    def class_rows(self):
        a = np.zeros(3, np.int32)
        _check(self._fn("get_class_rows")(self.h, a.ctypes.data_as(_ip)))
        return [int(v) for v in a]

    def shapes(self):
        """(option, threads, columns per thread) of each width class."""
        idx, thr, cpt = (np.zeros(3, np.int32) for _ in range(3))
        _check(self._fn("get_shapes")(self.h, idx.ctypes.data_as(_ip), thr.ctypes.data_as(_ip),
                                      cpt.ctypes.data_as(_ip)))
        return [(int(i), int(t), int(c)) for i, t, c in zip(idx, thr, cpt)]

    def carries(self):
        m = ctypes.c_int32()
        _check(self._fn("get_carries")(self.h, ctypes.byref(m)))
        return m.value

    def autotune(self, ticks, wrapX=False, wrapY=False, bc=None, cache_path=None):
        s, keep = make_bc(bc, 1, 1) if (bc is None or np.isscalar(bc)) else (None, ())
        r = ctypes.c_int32()
        _check(self._fn("autotune")(self.h, ticks, int(wrapX), int(wrapY), _bcp(s),
                                    None if cache_path is None else cache_path.encode(), ctypes.byref(r)))
        return r.value


class Adjoint2D(_Plan):
    """Per-cell-diffusivity ADI with adjoint gradients (zero flux, periodic
    or Dirichlet boundaries)."""
    _prefix = "adjoint2d"

    def __init__(self, ctx, X, Y):
        self.ctx, self.X, self.Y = ctx, X, Y
        self.h = _vp()
        _check(load().halhip_adjoint2d_create(ctx.h, X, Y, ctypes.byref(self.h)))

    def set_problem(self, diffusivity, decay, ticks, checkpoint_every, wrapX=False, wrapY=False, bc=None):
        """bc: None (zero flux), a constant Dirichlet value or edge arrays,
        on the non-wrapped axes."""
        D = _f64(diffusivity)
        _require(D.size == self.X * self.Y, "Adjoint2D.set_problem: diffusivity must hold xDim*yDim values")
        if not wrapX and not wrapY and bc is None:
            _check(self._fn("set_problem")(self.h, D.ctypes.data_as(_dp), decay, ticks, checkpoint_every))
            return
        s, keep = make_bc(bc, self.X, self.Y)
        _check(self._fn("set_problem_bc")(self.h, D.ctypes.data_as(_dp), decay, ticks, checkpoint_every,
                                          int(wrapX), int(wrapY), _bcp(s)))

    def forward(self, u0):
        u0 = _f32(u0)
        out = np.empty(self.X * self.Y, np.float32)
        _check(self._fn("forward")(self.h, u0.ctypes.data_as(_fp), out.ctypes.data_as(_fp)))
        return out

    def gradient(self, u0, target):
        """(J, dJ/dD float32 array, dJ/dk) for J = 1/2 |u_T - target|^2."""
        u0, target = _f32(u0), _f32(target)
        J = ctypes.c_double(0.0)
        gk = ctypes.c_double(0.0)
        gD = np.empty(self.X * self.Y, np.float32)
        _check(self._fn("gradient")(self.h, u0.ctypes.data_as(_fp), target.ctypes.data_as(_fp), ctypes.byref(J),
                                    gD.ctypes.data_as(_fp), ctypes.byref(gk)))
        return J.value, gD, gk.value
