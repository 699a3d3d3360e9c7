"""TF/SF plane-wave source on the device.

Reference: ``PlaneWaveSource`` (src/magnelio/sources/plane_wave.py):

* parameters and waveform           :63-198  (Gaussian amplitude *
  exp(-(t - t0)^2 / (2 sigma^2)), t0 = 4/f_max, sigma = 2/(pi f_max), or
  amplitude * sin(2 pi f t); the amplitude is also folded into the incident
  vectors, :249-252, so it enters squared — kept as the reference has it)
* box snapping                      :262-306
* face corrections                  :318-707
* injection times                   :710-734  (E at t_E - dt/2, H at t_H - dt/2)

The twelve E and twelve H face corrections are derived here from one rule
instead of being listed: on the face of the total-field box normal to axis
``a``, a tangential component ``c`` couples through the curl to the third
axis ``d``; with the Levi-Civita sign eps(c, a, d),

* E side (the E edge on the box face is total-field, the H face across it is
  scattered-field): min face adds ``-eps * H_inc_d``, max face
  ``+eps * H_inc_d``, times beta_E and the dual length along ``d``;
* H side (the H face just outside is scattered-field, the E edge on the box
  face total-field): min face (layer below the box) adds ``+eps * E_inc_d``,
  max face (layer at the box bound, only when a layer exists) ``-eps *
  E_inc_d``, times beta_H and the primal length along ``d``.

Faces run z, x, y (min then max), E components ascending and H components
descending within a face — the reference's order, which fixes the rounding
where two patches meet on a box edge.  Retardations are k.r / c0 at the
dual position of the layer across the face (E side) or at the face node
plane (H side).

Device form: per phase one correction list grouped by target element (CSR)
and a table of waveform samples per (step, distinct retardation), applied by
``mgh_tfsf_apply`` with the step index read on the device.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, layout
from .mesh import C0, ETA0

FACE_ORDER = ((2, 0), (2, 1), (0, 0), (0, 1), (1, 0), (1, 1))  # (axis, side): z, x, y
COMP = "xyz"


def _levi(a: int, b: int, c: int) -> int:
    return {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}[(a, b, c)]


def _dual_lengths(d: np.ndarray) -> np.ndarray:
    """Dual widths at nodes 0..N: d[0], mean of the two cells, d[N-1]."""
    n = d.size
    out = np.empty(n + 1)
    out[0] = d[0]
    if n > 1:
        out[1:n] = 0.5 * (d[: n - 1] + d[1:n])
    out[n] = d[n - 1]
    return out


def _classify_axis(k: np.ndarray) -> tuple[int, int]:
    ax = int(np.argmax(np.abs(k)))
    if abs(k[ax]) < 0.999:
        raise NotImplementedError(f"oblique incidence is not supported; direction must be axis-aligned, got {k}")
    return ax, int(np.sign(k[ax]))


@dataclass
class PlaneWaveSource:
    """Axis-aligned TF/SF plane wave (plane_wave.py:63-180 parameters)."""

    direction: tuple = (0.0, 0.0, 1.0)
    polarization: tuple = (1.0, 0.0, 0.0)
    corners: tuple | None = None
    amplitude: float = 1.0
    waveform: str = "gaussian"
    f_center: float | None = None
    f_max: float | None = None
    _attached: bool = field(default=False, repr=False, init=False)

    def __post_init__(self) -> None:
        d = np.array(self.direction, dtype=float)
        d /= np.linalg.norm(d)
        self.direction = tuple(d)
        p = np.array(self.polarization, dtype=float)
        p -= np.dot(p, d) * d
        norm_p = np.linalg.norm(p)
        if norm_p < 1e-10:
            raise ValueError("polarization must not be parallel to direction; "
                             f"got direction={self.direction}, polarization={self.polarization}")
        self.polarization = tuple(p / norm_p)

    # -- waveform (plane_wave.py:185-198) --------------------------------
    def excitation(self, t):
        if self.waveform == "gaussian":
            f_max = self.f_max or 1e9
            t0 = 4.0 / f_max
            sigma = 2.0 / (math.pi * f_max)
            return self.amplitude * np.exp(-((t - t0) ** 2) / (2 * sigma**2))
        if self.waveform == "sine":
            f = self.f_center or 1e9
            return self.amplitude * np.sin(2 * math.pi * f * t)
        raise ValueError(f"Unknown waveform: {self.waveform!r}")

    # -- geometry ---------------------------------------------------------
    def snap_box(self, grid) -> tuple[int, ...]:
        """(i0, i1, j0, j1, k0, k1): TF region = cells [i0, i1) x [j0, j1) x [k0, k1)."""
        n = grid.dims
        if self.corners is None:
            return (2, n[0] - 1, 2, n[1] - 1, 2, n[2] - 1)
        out = []
        for ax, nodes in enumerate((grid.x, grid.y, grid.z)):
            a, b = self.corners[0][ax], self.corners[1][ax]
            a = None if a is None or not np.isfinite(a) else float(a)
            b = None if b is None or not np.isfinite(b) else float(b)
            if a is not None and b is not None and b < a:
                a, b = b, a
            lo = 2 if a is None else int(np.searchsorted(nodes, a).clip(1, n[ax] - 1))
            hi = n[ax] - 1 if b is None else int(np.searchsorted(nodes, b).clip(lo + 1, n[ax]))
            out += [lo, hi]
        return tuple(out)

    def patches(self, grid, beta_e: np.ndarray, beta_h: np.ndarray, dtype):
        """Face corrections in the reference's order.

        Returns two lists (E, H) of (component axis, reference flat indices,
        coefficient array in ``dtype``, float64 retardation array, scalar),
        each array shaped like the patch; ``scalar`` marks the faces normal to
        the propagation axis, whose retardation the reference evaluates as one
        Python float (Python's ``**`` and numpy's array square differ in the
        last bit now and then, so the two paths are kept apart).  ``beta_e`` /
        ``beta_h`` are the solver's coefficients in the reference layout.
        """
        k = np.array(self.direction, dtype=float)
        prop, _ = _classify_axis(k)
        e_hat = np.array(self.polarization, dtype=float) * self.amplitude
        h_hat = np.cross(k, np.array(self.polarization, dtype=float)) * (self.amplitude / ETA0)
        box = self.snap_box(grid)
        lo, hi = box[0::2], box[1::2]
        nodes = (grid.x, grid.y, grid.z)
        ncell = grid.dims
        prim = tuple(np.diff(a) for a in nodes)
        dual = tuple(_dual_lengths(p) for p in prim)
        cent = tuple((a[:-1] + a[1:]) / 2 for a in nodes)
        nx, ny, nz = ncell
        shapes = layout.component_shapes(nx, ny, nz)
        e_off = {0: 0, 1: int(np.prod(shapes["Ex"])), 2: int(np.prod(shapes["Ex"]) + np.prod(shapes["Ey"]))}
        h_off = {0: 0, 1: int(np.prod(shapes["Hx"])), 2: int(np.prod(shapes["Hx"]) + np.prod(shapes["Hy"]))}

        def grids(ranges):
            return np.meshgrid(*[np.atleast_1d(r) for r in ranges], indexing="ij")

        def retard(pos):
            p = 0.0
            for kc, c in zip(k, pos):
                if kc != 0.0:
                    p = p + kc * c
            return np.asarray(p, dtype=np.float64) / C0

        out_e, out_h = [], []
        for a, side in FACE_ORDER:
            ia = lo[a] if side == 0 else hi[a]
            tang = [c for c in range(3) if c != a]
            # ---- E side: tangential components ascending
            for c in tang:
                d = 3 - a - c
                eps = _levi(c, a, d)
                factor = (-eps if side == 0 else eps) * h_hat[d]
                if factor == 0.0:
                    continue
                rng = [None] * 3
                rng[a] = np.array([ia])
                rng[c] = np.arange(lo[c], hi[c])
                rng[d] = np.arange(lo[d], hi[d] + 1)
                I, J, K = grids(rng)
                sh = shapes["E" + COMP[c]]
                flat = e_off[c] + (I * sh[1] + J) * sh[2] + K
                metric = dual[d][(I, J, K)[d]]
                coef = beta_e[flat].astype(dtype) * np.asarray(factor * metric, dtype=dtype)
                if side == 0:
                    pa = (nodes[a][ia - 1] + nodes[a][ia]) / 2
                else:
                    pa = (nodes[a][ia] + nodes[a][ia + 1]) / 2 if ia < ncell[a] else nodes[a][ia]
                pos = [None] * 3
                pos[a] = pa
                pos[c] = cent[c][(I, J, K)[c]]
                pos[d] = nodes[d][(I, J, K)[d]]
                out_e.append((c, flat, coef, np.broadcast_to(retard(pos), flat.shape), prop == a))
            # ---- H side: tangential components descending
            if side == 1 and ia >= ncell[a]:
                continue
            ih = ia - 1 if side == 0 else ia
            for c in reversed(tang):
                d = 3 - a - c
                eps = _levi(c, a, d)
                factor = (eps if side == 0 else -eps) * e_hat[d]
                if factor == 0.0:
                    continue
                rng = [None] * 3
                rng[a] = np.array([ih])
                rng[c] = np.arange(lo[c], hi[c] + 1)
                rng[d] = np.arange(lo[d], hi[d])
                I, J, K = grids(rng)
                sh = shapes["H" + COMP[c]]
                flat = h_off[c] + (I * sh[1] + J) * sh[2] + K
                metric = prim[d][(I, J, K)[d]]
                coef = beta_h[flat].astype(dtype) * np.asarray(factor * metric, dtype=dtype)
                pos = [None] * 3
                pos[a] = nodes[a][ia]
                pos[c] = nodes[c][(I, J, K)[c]]
                pos[d] = cent[d][(I, J, K)[d]]
                out_h.append((c, flat, coef, np.broadcast_to(retard(pos), flat.shape), prop == a))
        return out_e, out_h

    # -- device side ------------------------------------------------------
    def attach(self, solver) -> None:
        grid = solver.mesh.grid
        dims = solver.dims
        np_dt = solver._np_dtype
        dev = solver.e.device
        co = solver._coeffs
        pe, ph = self.patches(grid, np.asarray(co.beta_e), np.asarray(co.beta_h), np_dt)
        dt = solver.dt
        n_steps = solver.horizon
        n = np.arange(n_steps, dtype=np.float64) * dt
        t_e = (n + dt) - dt / 2                # inject_E(t + dt) at t_E - dt/2
        t_h = ((n + dt) + dt / 2) - dt / 2     # inject_H(t + dt + dt/2) at t_H - dt/2
        self._tables = []
        for patches, times, side in ((pe, t_e, "E"), (ph, t_h, "H")):
            self._tables.append(self._device_table(patches, times, side, dims, np_dt, dev))
        self.dtype_code = 4 if np_dt == np.float32 else 8
        self.nsets = solver.nsets
        self.n_steps = n_steps
        self._attached = True

    def _device_table(self, patches, times, side, dims, np_dt, dev):
        if not patches:
            return None
        flat = np.concatenate([p[1].ravel() for p in patches])
        coef = np.concatenate([p[2].ravel() for p in patches])
        delay = np.concatenate([p[3].ravel() for p in patches])
        scalar = np.concatenate([np.full(p[1].size, p[4]) for p in patches])
        keys = np.stack([delay, scalar.astype(np.float64)], axis=1)
        uniq, did = np.unique(keys, axis=0, return_inverse=True)
        did = did.ravel()
        # group by target, keeping patch order inside a target
        order = np.argsort(flat, kind="stable")
        flat_s, coef_s, did_s = flat[order], coef[order], did[order].astype(np.int32)
        targets, start = np.unique(flat_s, return_index=True)
        ptr = np.append(start, flat_s.size).astype(np.int32)
        dev_idx = layout.index_to_device(targets, *dims, side)
        wave = np.empty((times.size, uniq.shape[0]), dtype=np_dt)
        arr_cols = np.nonzero(uniq[:, 1] == 0.0)[0]
        if arr_cols.size:
            wave[:, arr_cols] = self.excitation(times[:, None] - uniq[arr_cols, 0][None, :]).astype(np_dt)
        for col in np.nonzero(uniq[:, 1] != 0.0)[0]:
            dly = float(uniq[col, 0])
            wave[:, col] = [self.excitation(float(t) - dly) for t in times]
        t = {
            "n_targets": int(targets.size),
            "targets": torch.from_numpy(dev_idx.astype(np.int64)).to(dev),
            "ptr": torch.from_numpy(ptr).to(dev),
            "coef": torch.from_numpy(np.ascontiguousarray(coef_s)).to(dev),
            "did": torch.from_numpy(np.ascontiguousarray(did_s)).to(dev),
            "wave": torch.from_numpy(np.ascontiguousarray(wave)).to(dev),
            "n_delays": int(uniq.shape[0]),
            "n_entries": int(flat.size),
        }
        return t

    def _apply(self, solver, stream, table, vec) -> None:
        if table is None:
            return
        _lib.call("mgh_tfsf_apply", self.dtype_code, self.nsets, vec.data_ptr(), vec.shape[1],
                  table["n_targets"], table["targets"].data_ptr(), table["ptr"].data_ptr(),
                  table["coef"].data_ptr(), table["did"].data_ptr(), table["wave"].data_ptr(),
                  table["n_delays"], self.n_steps, solver.step_ctr.data_ptr(), 0,
                  int(stream.cuda_stream))

    def inject_e(self, solver, stream) -> None:
        if self._attached:
            self._apply(solver, stream, self._tables[0], solver.e)

    def inject_h(self, solver, stream) -> None:
        if self._attached:
            self._apply(solver, stream, self._tables[1], solver.h)

    def device_bytes(self) -> int:
        tot = 0
        for t in self._tables:
            if t is not None:
                tot += sum(v.numel() * v.element_size() for v in t.values() if isinstance(v, torch.Tensor))
        return tot
