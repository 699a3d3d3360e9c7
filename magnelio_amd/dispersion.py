"""Pole-residue (ADE) materials on the device: eps(w) on the E edges and
mu(w) on the H faces.

Reference: ``DispersionOperator`` (src/magnelio/solver/_dispersion.py:160-569)
and its coefficient folding in ``FITTimeDomainSolver.setup``
(src/magnelio/solver/fit_td.py:382-403 for E, :430-461 for H).

A dispersive material is a :class:`magnelio_amd.mesh.Material` with
``epsilon`` = eps_inf and ``poles`` = ((a_p, r_p), ...) in rad/s, a real
``a_p`` giving a real pole and a complex one a conjugate pair stored once:
eps(w) = eps_inf + sum_p r_p / (j w - a_p) (+ conjugate partners); and/or
with ``mu`` = mu_inf and ``poles_mu`` for mu(w).

Host side (FP64, once): one block per dispersive material (sorted by id)
over the elements its cells own — the E edges by the clamped one-sided
lookup of M_sigma, PEC edges excluded, g = EPS0 A_dual / L_primal; the H
faces by the lookup of M_sigma_m, faces with M_mu = 0 excluded,
g = MU0 A_primal / L_dual — the trapezoidal pole coefficients
k = (1 + a dt/2)/(1 - a dt/2), c = r/(1 - a dt/2), and
W = g sum_p weight_p Re(c_p) whose half-step 0.5 dt W the solver folds into
alpha / beta of that side.  The H side is the same operator: the curl term
enters only as the right-hand side the kernel multiplies by beta, whatever
its sign, so the completion f -= beta j_hist and the pole advance are
identical (_dispersion.py:51-76).  Device side (csrc/mgh_ade.hip): the field
stash at the top of the step and the fused history subtraction + pole
advance right after that side's curl kernel, states in the field precision.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, layout
from .mesh import EPS0, MU0, _avg_d, _clamp_idx

TAB = 8


def _geom_e(grid) -> np.ndarray:
    ax, ay, az = _avg_d(grid.dx), _avg_d(grid.dy), _avg_d(grid.dz)
    dx, dy, dz = grid.dx, grid.dy, grid.dz
    gx = (ay[None, :, None] * az[None, None, :] / dx[:, None, None]).ravel()
    gy = (ax[:, None, None] * az[None, None, :] / dy[None, :, None]).ravel()
    gz = (ax[:, None, None] * ay[None, :, None] / dz[None, None, :]).ravel()
    return np.concatenate([gx, gy, gz])


def _geom_h(grid) -> np.ndarray:
    ax, ay, az = _avg_d(grid.dx), _avg_d(grid.dy), _avg_d(grid.dz)
    dx, dy, dz = grid.dx, grid.dy, grid.dz
    gx = (dy[None, :, None] * dz[None, None, :] / ax[:, None, None]).ravel()
    gy = (dx[:, None, None] * dz[None, None, :] / ay[None, :, None]).ravel()
    gz = (dx[:, None, None] * dy[None, :, None] / az[None, None, :]).ravel()
    return np.concatenate([gx, gy, gz])


def _face_owner(mat_id: np.ndarray) -> np.ndarray:
    nx, ny, nz = mat_id.shape
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    return np.concatenate([mat_id[il].ravel(), mat_id[:, jl].ravel(), mat_id[:, :, kl].ravel()])


def _edge_owner(mat_id: np.ndarray) -> np.ndarray:
    nx, ny, nz = mat_id.shape
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    return np.concatenate([mat_id[:, jl][:, :, kl].ravel(), mat_id[il][:, :, kl].ravel(),
                           mat_id[il][:, jl].ravel()])


class _Pole:
    def __init__(self, a, r, dt: float):
        k = (1.0 + a * dt / 2.0) / (1.0 - a * dt / 2.0)
        c = r / (1.0 - a * dt / 2.0)
        if complex(a).imag == 0.0:
            self.complex, self.weight = False, 1.0
            self.k, self.c = complex(k).real, complex(c).real
        else:
            self.complex, self.weight = True, 2.0
            self.k, self.c = complex(k), complex(c)

    def w_coeff(self) -> float:
        return self.weight * (self.c.real if self.complex else self.c)

    def row(self, dtype) -> np.ndarray:
        """kind, hw, 1+k, k, c (re/im) rounded like the reference's weak scalars."""
        opk = 1.0 + self.k
        t = np.zeros(TAB, dtype=dtype)
        t[0] = 1.0 if self.complex else 0.0
        t[1] = 0.5 * self.weight
        if self.complex:
            t[2], t[3] = opk.real, opk.imag
            t[4], t[5] = self.k.real, self.k.imag
            t[6], t[7] = self.c.real, self.c.imag
        else:
            t[2], t[4], t[6] = opk, self.k, self.c
        return t


class DeviceDispersion:
    """ADE state of every dispersive material of one side (one launch per hook)."""

    def __init__(self, blocks, n_f: int, dt: float, side: str = "E"):
        self.side = side
        self.blocks = blocks  # [(mat_id, element idx (reference layout), g float64, [poles])]
        W = np.zeros(n_f)
        for _, idx, g, poles in blocks:
            W[idx] += g * sum(p.w_coeff() for p in poles)
        self.W = W
        self.w_half = 0.5 * dt * W
        self._bound = False

    @classmethod
    def from_mesh(cls, mesh, dt: float, excluded: np.ndarray, side: str = "E"):
        """None when no material of ``side`` is dispersive (the plain path
        stays untouched).  ``excluded``: elements that carry no state — the
        unflattened PEC edge mask (E, _dispersion.py:296-306) or M_mu <= 0
        (H, the frozen faces, fit_td.py:436-441)."""
        if side not in ("E", "H"):
            raise ValueError(f"side must be 'E' or 'H', got {side!r}")
        attr = "poles" if side == "E" else "poles_mu"
        disp = [(mid, mat) for mid, mat in sorted(mesh.material_library.items())
                if getattr(mat, attr, ()) and not mat.is_pec]
        if not disp:
            return None
        if side == "E":
            own, const, geom = _edge_owner(mesh.material_id), EPS0, _geom_e(mesh.grid)
        else:
            own, const, geom = _face_owner(mesh.material_id), MU0, _geom_h(mesh.grid)
        excluded = np.asarray(excluded, dtype=bool)
        blocks = []
        for mid, mat in disp:
            idx = np.nonzero((own == mid) & ~excluded)[0]
            if idx.size == 0:
                continue
            blocks.append((mid, idx, const * geom[idx], [_Pole(a, r, dt) for a, r in getattr(mat, attr)]))
        if not blocks:
            return None
        return cls(blocks, own.size, dt, side)

    # -- device ---------------------------------------------------------
    def bind(self, solver) -> None:
        """Upload states once the final beta_E exists (beta sampled on the host)."""
        dims, K = solver.dims, solver.nsets
        np_dt = solver._np_dtype
        dev = solver.e.device
        beta = np.asarray(solver._coeffs.beta_e if self.side == "E" else solver._coeffs.beta_h)
        idx = np.concatenate([b[1] for b in self.blocks])
        self.n = int(idx.size)
        self.pmax = max(len(b[3]) for b in self.blocks)
        blk = np.concatenate([np.full(b[1].size, q, np.int32) for q, b in enumerate(self.blocks)])
        npoles = np.array([len(b[3]) for b in self.blocks], np.int32)
        tab = np.zeros((len(self.blocks), self.pmax, TAB), dtype=np_dt)
        for q, b in enumerate(self.blocks):
            for p, pole in enumerate(b[3]):
                tab[q, p] = pole.row(np_dt)
        g = np.concatenate([b[2] for b in self.blocks]).astype(np_dt)
        t_dt = solver.e.dtype

        def up(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)

        self.idx = up(layout.index_to_device(idx, *dims, self.side), torch.int64)
        self.g = up(g, t_dt)
        self.beta = up(beta[idx].astype(np_dt), t_dt)
        self.blk = up(blk, torch.int32)
        self.npoles = up(npoles, torch.int32)
        self.tab = up(tab, t_dt)
        self.fprev = torch.zeros((K, self.n), dtype=t_dt, device=dev)
        self.J = torch.zeros((K, self.pmax, self.n, 2), dtype=t_dt, device=dev)
        self.code = 4 if t_dt == torch.float32 else 8
        self.nsets = K
        self._bound = True

    def _field(self, solver):
        return solver.e if self.side == "E" else solver.h

    def save(self, solver, stream) -> None:
        f = self._field(solver)
        _lib.call("mgh_ade_save", self.code, self.nsets, f.data_ptr(), f.shape[1],
                  self.n, self.idx.data_ptr(), self.fprev.data_ptr(), int(stream.cuda_stream))

    def update(self, solver, stream) -> None:
        f = self._field(solver)
        _lib.call("mgh_ade_update", self.code, self.nsets, f.data_ptr(), f.shape[1],
                  self.n, self.idx.data_ptr(), self.g.data_ptr(), self.beta.data_ptr(),
                  self.fprev.data_ptr(), self.blk.data_ptr(), self.npoles.data_ptr(),
                  self.tab.data_ptr(), self.pmax, self.J.data_ptr(), int(stream.cuda_stream))

    def state_dict(self) -> dict:
        return {"J": self.J.cpu().numpy(), "f_prev": self.fprev.cpu().numpy()}

    def load_state_dict(self, sd: dict) -> None:
        self.J.copy_(torch.from_numpy(np.asarray(sd["J"])))
        self.fprev.copy_(torch.from_numpy(np.asarray(sd["f_prev"])))

    def reference_state(self, s: int) -> dict:
        """{str(mat_id): {"J0": ..., ...}} of field set s, the reference's
        DispersionOperator.state_dict (_dispersion.py:571-576): per block
        the pole currents over its edges, real for a real pole, complex for
        a stored-once conjugate pair (f_prev is rewritten before every use)."""
        J = self.J[s].cpu().numpy()
        out, off = {}, 0
        for mid, idx, _, poles in self.blocks:
            n = idx.size
            d = {}
            for i, p in enumerate(poles):
                re = J[i, off:off + n, 0]
                d[f"J{i}"] = (re + 1j * J[i, off:off + n, 1]) if p.complex else re.copy()
            out[str(mid)] = d
            off += n
        return out

    def load_reference_state(self, s: int, state: dict) -> None:
        J = self.J[s].cpu().numpy()
        off = 0
        for mid, idx, _, poles in self.blocks:
            n = idx.size
            d = state[str(mid)]
            for i, p in enumerate(poles):
                a = np.asarray(d[f"J{i}"])
                J[i, off:off + n, 0] = a.real
                J[i, off:off + n, 1] = a.imag if p.complex else 0.0
            off += n
        self.J[s].copy_(torch.from_numpy(J))
