"""Shared helpers for the test suite: golden-case decoding, error norms and
seeded inputs of the benchmark workloads."""
import numpy as np


def rel_l2(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(got - ref) / (den if den > 0 else 1.0)


def golden_cases(g):
    """Yields dicts for the golden diffusion / ADI cases."""
    for k in range(int(g["ncases"][0])):
        pre = f"case{k}_"
        X, Y, wx, wy, bck, adi = (int(v) for v in g[pre + "meta"])
        if bck == 0:
            bc = None
        elif bck == 1:
            bc = float(g[pre + "bc"][0])
        else:
            bc = tuple(g[pre + s] for s in ("bc_xlo", "bc_xhi", "bc_ylo", "bc_yhi"))
        yield dict(k=k, X=X, Y=Y, wrapX=wx, wrapY=wy, bc=bc, adi=bool(adi), rate=float(g[pre + "rate"][0]),
                   field=g[pre + "field"], deltas_in=g[pre + "deltas_in"], deltas_out=g[pre + "deltas_out"])


def advection_cases(g):
    """Yields dicts for the golden upwind-advection cases."""
    for k in range(int(g["nadv"][0])):
        pre = f"adv{k}_"
        X, Y, wx, wy, bck = (int(v) for v in g[pre + "meta"])
        bc = vbc = None
        if bck:
            bc = tuple(g[pre + s] for s in ("bc_xlo", "bc_xhi", "bc_ylo", "bc_yhi"))
            vbc = tuple(g[pre + s] for s in ("vbc_xlo", "vbc_xhi", "vbc_ylo", "vbc_yhi"))
        yield dict(k=k, X=X, Y=Y, wrapX=wx, wrapY=wy, bc=bc, vbc=vbc, field=g[pre + "field"],
                   deltas_in=g[pre + "deltas_in"], vel=tuple(float(v) for v in g[pre + "vel"]), xv=g[pre + "xv"],
                   yv=g[pre + "yv"], uniform_out=g[pre + "uniform_out"], cells_out=g[pre + "cells_out"])


def eigen_field(nx, ny):
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    return (np.sin(2 * np.pi * x / nx) * np.sin(2 * np.pi * y / ny)).ravel()


def lambda_cn(rate, n, k=1):
    s = np.sin(np.pi * k / n) ** 2
    return (1 - 2 * rate * s) / (1 + 2 * rate * s)


class RefGrid:
    """The reference's PDEGrid2D delta semantics in FP64 (host), with the
    oracle supplying the stencils' interface deltas."""

    def __init__(self, oracle, field, X, Y, wrap):
        self.o, self.u, self.X, self.Y, self.wrap = oracle, field.astype(np.float64), X, Y, wrap
        self.d = np.zeros(X * Y)

    def diffuse(self, rate, bc=None):
        self.d = self.o.diffuse(self.u, self.d, rate, self.X, self.Y, *self.wrap, bc=bc)

    def diffuse_rates(self, rx, ry):
        self.d = self.o.diffuse_rates(self.u, self.d, rx, ry, self.X, self.Y, *self.wrap)

    def advect(self, vx, vy, bc=None):
        self.d = self.o.advect(self.u, self.d, vx, vy, self.X, self.Y, *self.wrap, bc=bc)

    def advect_cells(self, xv, yv, bc=None, vbc=None):
        self.d = self.o.advect_cells(self.u, self.d, xv, yv, self.X, self.Y, *self.wrap, bc=bc, vbc=vbc)

    def adi(self, rate, bc=None):
        self.d = self.o.adi(self.u, self.d, rate, self.X, self.Y, *self.wrap, bc=bc)

    def point(self, op, i, v):
        if op == 0:  # Set
            self.d[i] = v - self.u[i]
        elif op == 1:  # Add
            self.d[i] += v
        elif op == 2:  # Mul
            self.d[i] += self.u[i] * v
        else:  # Scale
            self.d[i] += self.u[i] * (v - 1)

    def update(self):
        self.u = self.u + self.d
        self.d = np.zeros_like(self.d)
