"""Oracle: lumped RLC elements and discrete ports (CPU).  Test infrastructure only.

Restates (paths relative to the reference repository root):

* SeriesRLC / ParallelRLC   src/magnelio/circuit/companion.py:45-180 —
  trapezoidal companions V = r_eq I + v_hist; series: r_eq = R + 2L/dt +
  dt/(2C); parallel: Norton conductance G = 1/R + dt/(2L) + 2C/dt and
  v_hist = -I_hist / G
* edge-chain snapping       src/magnelio/ports/_lumped/factory.py:199-275
* LumpedElementOperator     src/magnelio/ports/_lumped/operator.py:134-166 —
  v_total = sum sign e (float64, chain order);
  i = ((v_src - v_hist) - v_total) / (r_eq + sum beta);
  e += sign beta i;  V = sum sign e (after), I = i
* source convention         operator.py:168-190 — v_src = 2 sqrt(Z0) a(t)
* beta                      factory.py:115-116 — dt / M_eps of the
  unflattened, lossless M_eps

Pinned by tests/test_oracle_lumped.py: the companion transients against
closed forms (step responses of series RL, series RC and parallel RC with
first-order convergence, the ringing frequency of an underdamped series
RLC — the reference's test_companion.py cases), and the Thévenin identity
V + Z0 I = v_src to 1e-12 in a field run (test_lumped_port.py:203-238).
"""

from __future__ import annotations

import math

import numpy as np

from . import yee


class SeriesRLC:
    def __init__(self, R=None, L=None, C=None):
        self.R, self.L, self.C = R, L, C
        self.i = self.vL = self.vC = 0.0

    def r_eq(self, dt):
        r = self.R if self.R is not None else 0.0
        if self.L is not None:
            r += 2.0 * self.L / dt
        if self.C is not None:
            r += dt / (2.0 * self.C)
        return r

    def v_hist(self, dt):
        vh = 0.0
        if self.L is not None:
            vh += -(2.0 * self.L / dt) * self.i - self.vL
        if self.C is not None:
            vh += self.vC + (dt / (2.0 * self.C)) * self.i
        return vh

    def advance(self, i, v, dt):
        if self.L is not None:
            self.vL = (2.0 * self.L / dt) * (i - self.i) - self.vL
        if self.C is not None:
            self.vC = self.vC + (dt / (2.0 * self.C)) * (i + self.i)
        self.i = i


class ParallelRLC:
    def __init__(self, R=None, L=None, C=None):
        self.R, self.L, self.C = R, L, C
        self.v = self.iL = self.iC = 0.0

    def g_eq(self, dt):
        g = 1.0 / self.R if self.R is not None else 0.0
        if self.L is not None:
            g += dt / (2.0 * self.L)
        if self.C is not None:
            g += 2.0 * self.C / dt
        return g

    def i_hist(self, dt):
        ih = 0.0
        if self.L is not None:
            ih += self.iL + (dt / (2.0 * self.L)) * self.v
        if self.C is not None:
            ih += -(2.0 * self.C / dt) * self.v - self.iC
        return ih

    def r_eq(self, dt):
        return 1.0 / self.g_eq(dt)

    def v_hist(self, dt):
        return -self.i_hist(dt) / self.g_eq(dt)

    def advance(self, i, v, dt):
        if self.L is not None:
            self.iL = self.iL + (dt / (2.0 * self.L)) * (v + self.v)
        if self.C is not None:
            self.iC = (2.0 * self.C / dt) * (v - self.v) - self.iC
        self.v = v


def snap_chain(x, y, z, start, end):
    """(axis, [(i, j, k)], [flat index in the reference E layout])."""
    nx, ny, nz = x.size - 1, y.size - 1, z.size - 1
    s = [int(np.argmin(np.abs(a - p))) for a, p in zip((x, y, z), start)]
    e = [int(np.argmin(np.abs(a - p))) for a, p in zip((x, y, z), end)]
    diff = [s[a] != e[a] for a in range(3)]
    if sum(diff) != 1:
        raise ValueError("start and end must differ along exactly one axis after snapping")
    ax = diff.index(True)
    lo, hi = sorted((s[ax], e[ax]))
    shp = yee.shapes(nx, ny, nz)
    comp = "E" + "xyz"[ax]
    off = {"Ex": 0, "Ey": int(np.prod(shp["Ex"])), "Ez": int(np.prod(shp["Ex"]) + np.prod(shp["Ey"]))}[comp]
    n0, n1, n2 = shp[comp]
    ijk, flat = [], []
    for m in range(lo, hi):
        p = list(s)
        p[ax] = m
        ijk.append(tuple(p))
        flat.append(off + (p[0] * n1 + p[1]) * n2 + p[2])
    return ax, ijk, flat


class Lumped:
    """A lumped element / discrete port on an edge chain; ``update_e`` and
    ``record`` follow the solver's port hooks."""

    def __init__(self, name, flat, beta, element, dt, Z0=0.0, signs=None, recorded=True):
        self.name = name
        self.dt = dt
        self.recorded = recorded
        self.flat = list(flat)
        self.beta = np.asarray(beta, dtype=float)
        self.signs = [1.0] * len(self.flat) if signs is None else list(signs)
        self.element = element
        self.Z0 = Z0
        self.beta_sum = float(np.sum(self.beta))
        self.wf = None
        self.last_I = 0.0
        self.V, self.I = [], []

    def set_excitation(self, fn):
        s = 2.0 * math.sqrt(self.Z0) if self.Z0 > 0.0 else 1.0
        self.wf = (lambda t, _f=fn, _s=s: _s * _f(t)) if s != 1.0 else fn

    def update_e(self, e, t):
        dt = self.dt
        v_total = 0.0
        for f, sg in zip(self.flat, self.signs):
            v_total += sg * float(e[f])
        v_src = self.wf(t) if self.wf is not None else 0.0
        r_eq = self.element.r_eq(dt)
        v_hist = self.element.v_hist(dt)
        i = (v_src - v_hist - v_total) / (r_eq + self.beta_sum)
        for f, sg, b in zip(self.flat, self.signs, self.beta):
            e[f] += sg * b * i
        self.element.advance(i, r_eq * i + v_hist, dt)
        self.last_I = i

    def record(self, e, h):
        v = 0.0
        for f, sg in zip(self.flat, self.signs):
            v += sg * float(e[f])
        self.V.append(v)
        self.I.append(self.last_I)
