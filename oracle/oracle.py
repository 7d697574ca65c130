"""Python handles on the CPU checkers -- TEST INFRASTRUCTURE ONLY.

``Oracle`` wraps ``oracle/liboracle.so`` (the FP64 restatement in
``hal_oracle.cpp``); ``Reference`` wraps ``oracle/_ref/libhalref.so`` (the
HAL-CPP reference CPU backends compiled from their own sources).  Only tests,
the smoke check and the CPU-baseline leg of the benchmark import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhalref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_up = ctypes.POINTER(ctypes.c_uint64)
_i32 = ctypes.c_int32
_f64 = ctypes.c_double


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _edges(bc, X, Y):
    """Edge arrays (bc(-1,y), bc(X,y), bc(x,-1), bc(x,Y)) from None, a
    constant or a 4-tuple of arrays."""
    if bc is None:
        return None, None, None, None
    if np.isscalar(bc):
        v = float(bc)
        return (np.full(Y, v), np.full(Y, v), np.full(X, v), np.full(X, v))
    xlo, xhi, ylo, yhi = bc
    return _f(xlo), _f(xhi), _f(ylo), _f(yhi)


class _Lib:
    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise FileNotFoundError(path + " is not built (run `make -C oracle` / `make -C oracle ref`)")
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def rng_doubles(self, seed, n):
        out = np.empty(n, np.float64)
        f = self.fn("rng_doubles")
        f.argtypes = [ctypes.c_uint64, ctypes.c_int64, _dp]
        f(seed, n, _p(out))
        return out

    def rng_save(self, seed, skip, path):
        """Hal::Rng(seed) after `skip` draws, written by Rng::Save."""
        f = self.fn("rng_save")
        f.restype = ctypes.c_int32
        f.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_char_p]
        assert f(seed, skip, path.encode()) == 0, path

    def rng_load_doubles(self, path, n):
        """n draws of a Hal::Rng restored by Rng::Load from `path`."""
        out = np.empty(n, np.float64)
        f = self.fn("rng_load_doubles")
        f.argtypes = [ctypes.c_char_p, ctypes.c_int64, _dp]
        f(path.encode(), n, _p(out))
        return out

    def rng_u64(self, seed, n):
        out = np.empty(n, np.uint64)
        f = self.fn("rng_u64")
        f.argtypes = [ctypes.c_uint64, ctypes.c_int64, _up]
        f(seed, n, _p(out, _up))
        return out

    def diffuse(self, field, deltas, rate, X, Y, wrapX, wrapY, bc=None):
        field = _f(field)
        deltas = _f(deltas).copy()
        e = _edges(bc, X, Y)
        f = self.fn("diffuse")
        f.argtypes = [_dp, _dp, _f64, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp]
        f(_p(field), _p(deltas), rate, X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e])
        return deltas

    def diffuse_rates(self, field, deltas, ratesX, ratesY, X, Y, wrapX, wrapY):
        field, ratesX, ratesY = _f(field), _f(ratesX), _f(ratesY)
        deltas = _f(deltas).copy()
        f = self.fn("diffuse_rates")
        f.argtypes = [_dp, _dp, _dp, _dp, _i32, _i32, _i32, _i32]
        f(_p(field), _p(deltas), _p(ratesX), _p(ratesY), X, Y, int(wrapX), int(wrapY))
        return deltas

    def adi(self, field, deltas, rate, X, Y, wrapX, wrapY, bc=None):
        field = _f(field)
        deltas = _f(deltas).copy()
        e = _edges(bc, X, Y)
        f = self.fn("adi")
        f.argtypes = [_dp, _dp, _f64, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp]
        f(_p(field), _p(deltas), rate, X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e])
        return deltas

    def agent_forces(self, pos, radii, R, k, X, Y, wrapX, wrapY):
        pos, radii = _f(pos), _f(radii)
        n = radii.shape[0]
        vel = np.zeros(2 * n, np.float64)
        f = self.fn("agent_forces")
        f.argtypes = [_dp, _dp, _i32, _dp, _f64, _f64, _i32, _i32, _i32, _i32]
        f(_p(pos), _p(radii), n, _p(vel), R, k, X, Y, int(wrapX), int(wrapY))
        return vel


    def advect(self, field, deltas, xVel, yVel, X, Y, wrapX, wrapY, bc=None):
        """deltas + uniform-velocity upwind advection (returns a new array)."""
        field = _f(field)
        deltas = _f(deltas).copy()
        e = _edges(bc, X, Y)
        f = self.fn("advect")
        f.argtypes = [_dp, _dp, _f64, _f64, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp]
        f(_p(field), _p(deltas), xVel, yVel, X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e])
        return deltas

    def advect_cells(self, field, deltas, xVels, yVels, X, Y, wrapX, wrapY, bc=None, vbc=None):
        """deltas + per-cell-velocity upwind advection; bc / vbc: field and
        velocity edge values (both or neither)."""
        field, xVels, yVels = _f(field), _f(xVels), _f(yVels)
        deltas = _f(deltas).copy()
        e = _edges(bc, X, Y)
        v = _edges(vbc, X, Y)
        f = self.fn("advect_cells")
        f.argtypes = [_dp, _dp, _dp, _dp, _i32, _i32, _i32, _i32] + [_dp] * 8
        f(_p(field), _p(deltas), _p(xVels), _p(yVels), X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e],
          *[_p(a) for a in v])
        return deltas


class Oracle(_Lib):
    """FP64 restatement (hal_oracle.cpp)."""

    def __init__(self, path=ORACLE_SO):
        super().__init__(path, "oracle_")

    def set_threads(self, n):
        """Worker threads of the line loops (ADI and the per-cell ADI and its
        adjoint); the results do not depend on the count."""
        f = self.fn("set_threads")
        f.argtypes = [_i32]
        f(int(n))

    def model_ticks(self, field, ticks, rate, xVel, yVel, decay, X, Y, wrapX, wrapY, bc=None, src_idx=(),
                    src_val=()):
        """Diffusion + Advection + MulAll + Add + Update ticks (FP64)."""
        field = _f(field).copy()
        e = _edges(bc, X, Y)
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("model_ticks")
        f.argtypes = [_dp, _i32, _f64, _f64, _f64, _f64, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp, _i32, _lp, _dp]
        f(_p(field), ticks, rate, xVel, yVel, decay, X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e], len(si),
          _p(si, _lp), _p(sv))
        return field

    def explicit_ticks(self, field, ticks, rate, decay, X, Y, wrapX, wrapY, src_idx=(), src_val=()):
        field = _f(field).copy()
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("explicit_ticks")
        f.argtypes = [_dp, _i32, _f64, _f64, _i32, _i32, _i32, _i32, _i32, _lp, _dp]
        f(_p(field), ticks, rate, decay, X, Y, int(wrapX), int(wrapY), len(si), _p(si, _lp), _p(sv))
        return field

    def adi_ticks(self, field, ticks, rate, decay, X, Y, wrapX, wrapY, bc=None, src_idx=(), src_val=()):
        field = _f(field).copy()
        e = _edges(bc, X, Y)
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("adi_ticks")
        f.argtypes = [_dp, _i32, _f64, _f64, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp, _i32, _lp, _dp]
        f(_p(field), ticks, rate, decay, X, Y, int(wrapX), int(wrapY), *[_p(a) for a in e],
          len(si), _p(si, _lp), _p(sv))
        return field

    def thomas(self, a, b, c, d, periodic=False):
        a, b, c, d = _f(a), _f(b), _f(c), _f(d)
        x = np.empty_like(d)
        f = self.fn("thomas_periodic" if periodic else "thomas")
        f.argtypes = [_i32, _dp, _dp, _dp, _dp, _dp]
        f(len(d), _p(a), _p(b), _p(c), _p(d), _p(x))
        return x

    def adi_line(self, d, rate, wrap, has_bc):
        d = _f(d)
        x = np.empty_like(d)
        f = self.fn("adi_line")
        f.argtypes = [_i32, _f64, _i32, _i32, _dp, _dp]
        f(len(d), rate, int(wrap), int(has_bc), _p(d), _p(x))
        return x

    def agent_bins(self, pos, X, Y):
        pos = _f(pos)
        n = pos.shape[0] // 2
        offsets = np.zeros(X * Y + 1, np.int32)
        bucket = np.zeros(max(n, 1), np.int32)
        f = self.fn("agent_bins")
        f.argtypes = [_dp, _i32, _i32, _i32, _ip, _ip]
        f(_p(pos), n, X, Y, _p(offsets, _ip), _p(bucket, _ip))
        return offsets, bucket[:n]

    def agent_move(self, pos, vel, dt, X, Y, wrapX, wrapY):
        pos = _f(pos).copy()
        vel = _f(vel)
        f = self.fn("agent_move")
        f.argtypes = [_dp, _dp, _i32, _f64, _i32, _i32, _i32, _i32]
        f(_p(pos), _p(vel), pos.shape[0] // 2, dt, X, Y, int(wrapX), int(wrapY))
        return pos

    def agent_ticks_motion(self, pos, carry, radii, ticks, R, k, dt, friction, max_vel, X, Y, wrapX, wrapY):
        """Returns (positions, carried velocities) after `ticks` ticks."""
        pos = _f(pos).copy()
        carry = _f(carry).copy()
        radii = _f(radii)
        f = self.fn("agent_ticks_motion")
        f.argtypes = [_dp, _dp, _dp, _i32, _i32, _f64, _f64, _f64, _f64, _f64, _i32, _i32, _i32, _i32]
        f(_p(pos), _p(carry), _p(radii), radii.shape[0], ticks, R, k, dt, friction, max_vel, X, Y, int(wrapX),
          int(wrapY))
        return pos, carry

    def agent_ticks(self, pos, radii, ticks, R, k, dt, X, Y, wrapX, wrapY):
        pos = _f(pos).copy()
        radii = _f(radii)
        f = self.fn("agent_ticks")
        f.argtypes = [_dp, _dp, _i32, _i32, _f64, _f64, _f64, _i32, _i32, _i32, _i32]
        f(_p(pos), _p(radii), radii.shape[0], ticks, R, k, dt, X, Y, int(wrapX), int(wrapY))
        return pos


    @staticmethod
    def cell_edges(bc, X, Y):
        """Ghost values of the per-cell model as one [xlo(Y) | xhi(Y) | ylo(X) |
        yhi(X)] array: None (zero flux), a constant, or a 4-tuple of edge arrays."""
        if bc is None:
            return None
        if np.isscalar(bc):
            return np.full(2 * X + 2 * Y, float(bc))
        e = np.concatenate([_f(a) for a in bc])
        assert e.size == 2 * X + 2 * Y
        return e

    def adi_cell_forward(self, u0, D, k, ticks, X, Y, wrapX=False, wrapY=False, bc=None):
        u = _f(u0).copy()
        D = _f(D)
        e = self.cell_edges(bc, X, Y)
        f = self.fn("adi_cell_forward_bc")
        f.argtypes = [_dp, _dp, _f64, _i32, _i32, _i32, _i32, _i32, _dp]
        f(_p(u), _p(D), k, ticks, X, Y, int(wrapX), int(wrapY), None if e is None else _p(e))
        return u

    def adi_cell_explicit(self, field, D, X, Y, axis, wrapX=False, wrapY=False, bc=None, with_bc=True):
        """field + 1/2 L field along one axis (0: x, 1: y)."""
        field, D = _f(field), _f(D)
        out = np.empty_like(field)
        e = self.cell_edges(bc, X, Y)
        f = self.fn("adi_cell_explicit")
        f.argtypes = [_dp, _dp, _dp, _i32, _i32, _i32, _i32, _i32, _dp, _i32]
        f(_p(field), _p(out), _p(D), X, Y, axis, int(wrapX), int(wrapY), None if e is None else _p(e),
          int(with_bc))
        return out

    def adi_cell_tick_transpose(self, lam, D, k, X, Y, wrapX=False, wrapY=False, bc=None):
        lam = _f(lam)
        D = _f(D)
        out = np.empty_like(lam)
        e = self.cell_edges(bc, X, Y)
        f = self.fn("adi_cell_tick_transpose_bc")
        f.argtypes = [_dp, _dp, _dp, _f64, _i32, _i32, _i32, _i32, _dp]
        f(_p(lam), _p(out), _p(D), k, X, Y, int(wrapX), int(wrapY), None if e is None else _p(e))
        return out

    def adi_cell_gradient(self, u0, D, k, ticks, X, Y, target, wrapX=False, wrapY=False, bc=None):
        """(J, dJ/dD (X*Y), dJ/dk) for J = 1/2 |u_T - target|^2."""
        u0, D, target = _f(u0), _f(D), _f(target)
        gD = np.zeros(X * Y)
        gk = ctypes.c_double(0.0)
        e = self.cell_edges(bc, X, Y)
        f = self.fn("adi_cell_gradient_bc")
        f.restype = ctypes.c_double
        f.argtypes = [_dp, _dp, _f64, _i32, _i32, _i32, _dp, _dp, ctypes.POINTER(ctypes.c_double), _i32, _i32,
                      _dp]
        J = f(_p(u0), _p(D), k, ticks, X, Y, _p(target), _p(gD), ctypes.byref(gk), int(wrapX), int(wrapY),
              None if e is None else _p(e))
        return J, gD, gk.value


class Reference(_Lib):
    """The compiled HAL-CPP reference (oracle/_ref/libhalref.so)."""

    def __init__(self, path=REF_SO):
        super().__init__(path, "ref_")

    def grid_ticks(self, field, ticks, adi, rate, decay, X, Y, wrapX, wrapY, bc_value=None,
                   src_idx=(), src_val=()):
        field = _f(field).copy()
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("grid_ticks")
        f.argtypes = [_dp, _i32, _i32, _f64, _f64, _i32, _i32, _i32, _i32, _i32, _f64, _i32, _lp, _dp]
        f(_p(field), ticks, int(adi), rate, decay, X, Y, int(wrapX), int(wrapY),
          0 if bc_value is None else 1, 0.0 if bc_value is None else float(bc_value),
          len(si), _p(si, _lp), _p(sv))
        return field


    def model_ticks(self, field, ticks, rate, xVel, yVel, decay, X, Y, wrapX, wrapY, bc_value=None, src_idx=(),
                    src_val=()):
        """PDEGrid2D Diffusion + Advection + MulAll + Add + Update ticks."""
        field = _f(field).copy()
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("model_ticks")
        f.argtypes = [_dp, _i32, _f64, _f64, _f64, _f64, _i32, _i32, _i32, _i32, _i32, _f64, _i32, _lp, _dp]
        f(_p(field), ticks, rate, xVel, yVel, decay, X, Y, int(wrapX), int(wrapY), 0 if bc_value is None else 1,
          0.0 if bc_value is None else float(bc_value), len(si), _p(si, _lp), _p(sv))
        return field

    def save_pdegrid2d(self, path, field, deltas, X, Y, wrapX, wrapY, update_ct):
        """Hal::SavePDEGrid2D of a grid holding field / deltas."""
        f = self.fn("save_pdegrid2d")
        f.restype = ctypes.c_int32
        f.argtypes = [ctypes.c_char_p, _dp, _dp, _i32, _i32, _i32, _i32, _i32]
        rc = f(path.encode(), _p(_f(field)), _p(_f(deltas)), X, Y, int(wrapX), int(wrapY), update_ct)
        assert rc == 0, path

    def load_pdegrid2d(self, path, capacity):
        """Hal::LoadPDEGrid2D: (X, Y, wrapX, wrapY, updateCt, field, deltas) or None when it throws."""
        dims = np.zeros(5, np.int32)
        field = np.zeros(capacity)
        deltas = np.zeros(capacity)
        f = self.fn("load_pdegrid2d")
        f.restype = ctypes.c_int32
        f.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), _dp, _dp, ctypes.c_int64]
        if f(path.encode(), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _p(field), _p(deltas),
             capacity) != 0:
            return None
        n = int(dims[0]) * int(dims[1])
        return (*(int(v) for v in dims), field[:n], deltas[:n])

    def grid_ticks_timed(self, field, warmup, ticks, adi, rate, decay, X, Y, wrapX, wrapY, bc_value=None,
                         src_idx=(), src_val=()):
        """Seconds for `ticks` PDEGrid2D model ticks of the compiled reference."""
        field = _f(field).copy()
        si = np.ascontiguousarray(src_idx, np.int64)
        sv = _f(src_val)
        f = self.fn("grid_ticks_timed")
        f.restype = ctypes.c_double
        f.argtypes = [_dp, _i32, _i32, _i32, _f64, _f64, _i32, _i32, _i32, _i32, _i32, _f64, _i32, _lp, _dp]
        secs = f(_p(field), warmup, ticks, int(adi), rate, decay, X, Y, int(wrapX), int(wrapY),
                 0 if bc_value is None else 1, 0.0 if bc_value is None else float(bc_value),
                 len(si), _p(si, _lp), _p(sv))
        return secs, field

    def agent_forces_timed(self, pos, radii, calls, R, k, X, Y, wrapX, wrapY):
        """Seconds for `calls` CpuAgentForceBackend2D::AccumulateForces calls."""
        pos, radii = _f(pos), _f(radii)
        f = self.fn("agent_forces_timed")
        f.restype = ctypes.c_double
        f.argtypes = [_dp, _dp, _i32, _i32, _f64, _f64, _i32, _i32, _i32, _i32]
        return f(_p(pos), _p(radii), radii.size, calls, R, k, X, Y, int(wrapX), int(wrapY))

    def diffuse_calls_timed(self, field, rate, X, Y, wrapX, wrapY, adi, calls):
        """Seconds for `calls` CpuDiffusionBackend2D::Diffuse (or
        CpuAdiDiffusionBackend2D::DiffuseADI) calls after one untimed call."""
        field = _f(field)
        f = self.fn("diffuse_calls_timed")
        f.restype = ctypes.c_double
        f.argtypes = [_dp, _f64, _i32, _i32, _i32, _i32, _i32, _i32]
        return f(_p(field), rate, X, Y, int(wrapX), int(wrapY), int(adi), calls)

    def adi_ticks_timed(self, field, warmup, ticks, rate, X, Y, bc_value):
        """Seconds for `ticks` Dirichlet ADI ticks of the compiled reference."""
        field = _f(field).copy()
        f = self.fn("adi_ticks_timed")
        f.restype = ctypes.c_double
        f.argtypes = [_dp, _i32, _i32, _f64, _i32, _i32, _f64]
        secs = f(_p(field), warmup, ticks, rate, X, Y, bc_value)
        return secs, field


def available_reference():
    return os.path.exists(REF_SO)
