"""Lumped RLC elements and discrete (lumped) ports on the device.

Reference counterparts (relative to the reference repository root):

* ``SeriesRLC`` / ``ParallelRLC``        src/magnelio/circuit/companion.py:45-180
* ``PortSpecLumped`` / ``build_lumped_port`` / ``build_lumped_element``
                                          src/magnelio/ports/_lumped/factory.py:33-197
* edge-chain snapping                    factory.py:199-275
* ``LumpedElementOperator.update_e``     src/magnelio/ports/_lumped/operator.py:134-166
* source convention v_src = 2 sqrt(Z0) a(t)   operator.py:168-190

The companion parameters (r_eq and the trapezoidal gains) are constants of
a run and are folded on the host in the reference's operation order; the
per-step recursion, the edge gather/scatter and the V/I recording run on
the device (csrc/mgh_lumped.hip).  beta = dt / M_eps of the unflattened,
lossless M_eps (factory.py:115-116).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, layout


def _validate_rlc(R, L, C, kind: str) -> None:
    if R is None and L is None and C is None:
        raise ValueError(f"{kind}: at least one of R, L, C must be given.")
    for name, val in (("R", R), ("L", L), ("C", C)):
        if val is not None and not (val > 0.0):
            raise ValueError(f"{kind}: {name} must be positive; got {val}.")


@dataclass
class SeriesRLC:
    """Series R-L-C companion (shared current); any subset present."""

    R: float | None = None
    L: float | None = None
    C: float | None = None

    def __post_init__(self) -> None:
        _validate_rlc(self.R, self.L, self.C, "SeriesRLC")

    def r_eq(self, dt: float) -> float:
        r = self.R if self.R is not None else 0.0
        if self.L is not None:
            r += 2.0 * self.L / dt
        if self.C is not None:
            r += dt / (2.0 * self.C)
        return r

    def device_params(self, dt: float, beta_sum: float) -> np.ndarray:
        kl = 2.0 * self.L / dt if self.L is not None else 0.0
        kc = dt / (2.0 * self.C) if self.C is not None else 0.0
        return np.array([0.0, self.L is not None, self.C is not None, self.r_eq(dt), beta_sum,
                         kl, kc, 0.0], dtype=np.float64)


@dataclass
class ParallelRLC:
    """Parallel R-L-C companion (shared voltage); any subset present."""

    R: float | None = None
    L: float | None = None
    C: float | None = None

    def __post_init__(self) -> None:
        _validate_rlc(self.R, self.L, self.C, "ParallelRLC")

    def g_eq(self, dt: float) -> float:
        g = 1.0 / self.R if self.R is not None else 0.0
        if self.L is not None:
            g += dt / (2.0 * self.L)
        if self.C is not None:
            g += 2.0 * self.C / dt
        return g

    def r_eq(self, dt: float) -> float:
        return 1.0 / self.g_eq(dt)

    def device_params(self, dt: float, beta_sum: float) -> np.ndarray:
        kl = dt / (2.0 * self.L) if self.L is not None else 0.0
        kc = 2.0 * self.C / dt if self.C is not None else 0.0
        return np.array([1.0, self.L is not None, self.C is not None, self.r_eq(dt), beta_sum,
                         kl, kc, self.g_eq(dt)], dtype=np.float64)


@dataclass
class PortSpecLumped:
    """Discrete port between two points (factory.py:33-64)."""

    name: str
    start: tuple
    end: tuple
    Z0: float = 50.0
    element: SeriesRLC | ParallelRLC | None = None


@dataclass
class LumpedElement:
    """Passive lumped element between two points (no excitation, Z0 = 0)."""

    name: str
    start: tuple
    end: tuple
    element: SeriesRLC | ParallelRLC


def snap_edge_chain(grid, start, end):
    """(direction, reference flat E indices, edge lengths) — factory.py:199-275."""
    nx, ny, nz = grid.dims
    s = [int(np.argmin(np.abs(a - p))) for a, p in zip((grid.x, grid.y, grid.z), start)]
    e = [int(np.argmin(np.abs(a - p))) for a, p in zip((grid.x, grid.y, grid.z), end)]
    diff = [s[a] != e[a] for a in range(3)]
    if sum(diff) != 1:
        raise ValueError("start and end must differ along exactly one Cartesian axis after grid "
                         f"snapping; got start node {tuple(s)}, end node {tuple(e)}")
    ax = diff.index(True)
    lo, hi = sorted((s[ax], e[ax]))
    shapes = layout.component_shapes(nx, ny, nz)
    comp = "E" + "xyz"[ax]
    off = {"Ex": 0, "Ey": int(np.prod(shapes["Ex"])),
           "Ez": int(np.prod(shapes["Ex"]) + np.prod(shapes["Ey"]))}[comp]
    n0, n1, n2 = shapes[comp]
    limits = (nx, ny, nz)
    d = (grid.dx, grid.dy, grid.dz)[ax]
    flat, dl = [], []
    for m in range(lo, hi):
        p = list(s)
        p[ax] = m
        # clamp like the reference builder
        p = [max(0, min(p[a], limits[a] - 1 if a == ax else limits[a])) for a in range(3)]
        flat.append(off + (p[0] * n1 + p[1]) * n2 + p[2])
        dl.append(float(d[p[ax]]))
    if sum(dl) == 0.0:
        raise ValueError("zero-length edge chain: start and end coincide on the snapped grid")
    return "xyz"[ax], flat, dl


class LumpedOperator:
    """A lumped element or discrete port resolved onto the mesh."""

    n_modes = 1

    def __init__(self, name, flat, beta, element, Z0=0.0, signs=None, recorded=True):
        self.name = name
        self.flat = np.asarray(flat, dtype=np.int64)
        self.beta = np.asarray(beta, dtype=np.float64)
        self.signs = np.ones(self.flat.size) if signs is None else np.asarray(signs, dtype=np.float64)
        self.element = element
        self.Z0 = float(Z0)
        self.recorded = recorded
        # numpy's sum, as the reference folds it (operator.py:130)
        self.beta_sum = float(np.sum(self.beta))

    def source_scale(self) -> float:
        return 2.0 * math.sqrt(self.Z0) if self.Z0 > 0.0 else 1.0


def build_lumped_port(spec: PortSpecLumped, mesh, m_eps: np.ndarray, dt: float) -> LumpedOperator:
    """factory.py:67-137 (no symmetry planes in this build)."""
    _, flat, _ = snap_edge_chain(mesh.grid, spec.start, spec.end)
    beta = dt / np.asarray(m_eps[flat], dtype=float)
    elem = spec.element if spec.element is not None else SeriesRLC(R=spec.Z0)
    return LumpedOperator(spec.name, flat, beta, elem, Z0=spec.Z0, recorded=True)


def build_lumped_element(spec: LumpedElement, mesh, m_eps: np.ndarray, dt: float) -> LumpedOperator:
    """factory.py:140-196: a passive load, never excited, not recorded."""
    _, flat, _ = snap_edge_chain(mesh.grid, spec.start, spec.end)
    beta = dt / np.asarray(m_eps[flat], dtype=float)
    return LumpedOperator(spec.name, flat, beta, spec.element, Z0=0.0, recorded=False)


class DeviceLumped:
    """Device state, tables and the descriptor of every lumped operator."""

    WORDS = 12

    def __init__(self, ops, solver):
        self.ops = list(ops)
        dims, K = solver.dims, solver.nsets
        dev = solver.e.device
        dt = solver.dt
        n_steps = solver.horizon
        seen = set()
        for op in self.ops:
            edges = set(op.flat.tolist())
            if edges & seen:
                raise ValueError(f"lumped operator {op.name!r} shares edges with another one")
            seen |= edges
        names = [op.name for op in self.ops]
        if len(set(names)) != len(names):
            raise ValueError("lumped operator names must be unique")
        self.desc = np.zeros(self.WORDS * len(self.ops), dtype=np.int64)
        self._keep = []
        self.state, self.V_rec, self.I_rec = [], [], []
        t = np.arange(n_steps, dtype=np.float64) * dt + dt  # update_e(fields, t + dt)
        edge_lists = []
        for q, op in enumerate(self.ops):
            idx = layout.index_to_device(op.flat, *dims, "E")
            edge_lists.append(idx)
            exc = np.zeros(K, dtype=np.int32)
            src = np.zeros((K, n_steps), dtype=np.float64)
            for s, ch in enumerate(solver.excitations):
                if ch is not None and ch.port == op.name:
                    if not op.recorded:
                        raise ValueError(f"{op.name!r} is a passive element and cannot be excited")
                    if ch.mode != 0:
                        raise ValueError(f"lumped port {op.name!r} has only mode 0")
                    exc[s] = 1
                    scale = op.source_scale()
                    wf = ch.waveform
                    if scale != 1.0:
                        src[s] = [scale * wf(float(tt)) for tt in t]
                    else:
                        src[s] = [wf(float(tt)) for tt in t]

            def dev_t(a, dtype=torch.float64):
                ten = torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)
                self._keep.append(ten)
                return ten

            t_idx = dev_t(idx, torch.int64)
            t_sign = dev_t(op.signs)
            t_beta = dev_t(op.beta)
            t_prm = dev_t(op.element.device_params(dt, op.beta_sum))
            t_state = dev_t(np.zeros((K, 5)))
            t_exc = dev_t(exc, torch.int32)
            t_src = dev_t(src)
            self.state.append(t_state)
            if op.recorded:
                vr = dev_t(np.zeros((K, n_steps)))
                ir = dev_t(np.zeros((K, n_steps)))
            else:
                vr = ir = None
            self.V_rec.append(vr)
            self.I_rec.append(ir)
            w = self.desc[q * self.WORDS:(q + 1) * self.WORDS]
            w[0] = op.flat.size
            w[1], w[2], w[3] = t_idx.data_ptr(), t_sign.data_ptr(), t_beta.data_ptr()
            w[4], w[5], w[6], w[7] = (t_prm.data_ptr(), t_state.data_ptr(), t_exc.data_ptr(),
                                      t_src.data_ptr())
            w[8] = n_steps
            w[9] = vr.data_ptr() if vr is not None else 0
            w[10] = ir.data_ptr() if ir is not None else 0
            w[11] = n_steps if vr is not None else 0
        self.dtype_code = 4 if solver.e.dtype == torch.float32 else 8
        # device-layout E edges the elements write (kept live by dead-unit skipping)
        self.edges = np.concatenate(edge_lists) if edge_lists else np.zeros(0, np.int64)

    def step(self, solver, stream) -> None:
        if not self.ops:
            return
        _lib.call("mgh_lumped_step", self.dtype_code, solver.nsets, solver.e.data_ptr(),
                  solver.e.shape[1], int(self.desc.ctypes.data), len(self.ops),
                  solver.step_ctr.data_ptr(), int(stream.cuda_stream))

    def recorded(self, s: int, n: int) -> dict:
        out = {}
        for op, V, I in zip(self.ops, self.V_rec, self.I_rec):
            if V is not None:
                out[(op.name, 0)] = (V[s, :n].cpu().numpy().copy(), I[s, :n].cpu().numpy().copy())
        return out

    def state_dict(self) -> dict:
        sd = {}
        for op, st, V, I in zip(self.ops, self.state, self.V_rec, self.I_rec):
            sd[f"{op.name}_state"] = st.cpu().numpy()
            if V is not None:
                sd[f"{op.name}_V"] = V.cpu().numpy()
                sd[f"{op.name}_I"] = I.cpu().numpy()
        return sd

    def set_state(self, s: int, n: int) -> dict:
        """Companion state and V/I records of field set s (first n steps)."""
        sd = {}
        for op, st, V, I in zip(self.ops, self.state, self.V_rec, self.I_rec):
            d = {"state": st[s].cpu().numpy()}
            if V is not None:
                d["V"] = V[s, :n].cpu().numpy()
                d["I"] = I[s, :n].cpu().numpy()
            sd[op.name] = d
        return sd

    def load_set_state(self, s: int, sd: dict, n: int) -> None:
        for op, st, V, I in zip(self.ops, self.state, self.V_rec, self.I_rec):
            d = sd[op.name]
            st[s].copy_(torch.from_numpy(np.asarray(d["state"])))
            if V is not None and n:
                V[s, :n].copy_(torch.from_numpy(np.asarray(d["V"])))
                I[s, :n].copy_(torch.from_numpy(np.asarray(d["I"])))

    def load_state_dict(self, sd: dict) -> None:
        for op, st, V, I in zip(self.ops, self.state, self.V_rec, self.I_rec):
            st.copy_(torch.from_numpy(np.asarray(sd[f"{op.name}_state"])))
            if V is not None:
                V.copy_(torch.from_numpy(np.asarray(sd[f"{op.name}_V"])))
                I.copy_(torch.from_numpy(np.asarray(sd[f"{op.name}_I"])))
