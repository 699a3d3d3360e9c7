"""Pole-residue (ADE) dielectrics on the device, E side.

Reference: ``DispersionOperator`` (src/magnelio/solver/_dispersion.py:160-533)
and its coefficient folding in ``FITTimeDomainSolver.setup``
(src/magnelio/solver/fit_td.py:382-403).

A dispersive material is a :class:`magnelio_amd.mesh.Material` with
``epsilon`` = eps_inf and ``poles`` = ((a_p, r_p), ...) in rad/s, a real
``a_p`` giving a real pole and a complex one a conjugate pair stored once:
eps(w) = eps_inf + sum_p r_p / (j w - a_p) (+ conjugate partners).

Host side (FP64, once): one block per dispersive material (sorted by id)
over the E edges its cells own (the clamped one-sided lookup of the mass
matrices, PEC edges excluded), g = EPS0 A_dual / L_primal, the trapezoidal
pole coefficients k = (1 + a dt/2)/(1 - a dt/2), c = r/(1 - a dt/2), and
W = g sum_p weight_p Re(c_p) whose half-step 0.5 dt W the solver folds into
alpha_E / beta_E.  Device side (csrc/mgh_ade.hip): the field stash at the
top of the step and the fused history subtraction + pole advance right
after the curl kernel, states in the field precision.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, layout
from .mesh import EPS0, _avg_d, _clamp_idx

TAB = 8


def _geom_e(grid) -> np.ndarray:
    ax, ay, az = _avg_d(grid.dx), _avg_d(grid.dy), _avg_d(grid.dz)
    dx, dy, dz = grid.dx, grid.dy, grid.dz
    gx = (ay[None, :, None] * az[None, None, :] / dx[:, None, None]).ravel()
    gy = (ax[:, None, None] * az[None, None, :] / dy[None, :, None]).ravel()
    gz = (ax[:, None, None] * ay[None, :, None] / dz[None, None, :]).ravel()
    return np.concatenate([gx, gy, gz])


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
    """E-side ADE state of every dispersive material (one launch per hook)."""

    def __init__(self, blocks, n_e: int, dt: float):
        self.blocks = blocks  # [(mat_id, edge idx (reference layout), g float64, [poles])]
        W = np.zeros(n_e)
        for _, idx, g, poles in blocks:
            W[idx] += g * sum(p.w_coeff() for p in poles)
        self.W = W
        self.w_half = 0.5 * dt * W
        self._bound = False

    @classmethod
    def from_mesh(cls, mesh, dt: float, pec_e_flat: np.ndarray):
        """None when no material is dispersive (the plain path stays untouched)."""
        disp = [(mid, mat) for mid, mat in sorted(mesh.material_library.items())
                if getattr(mat, "poles", ()) and not mat.is_pec]
        if not disp:
            return None
        own = _edge_owner(mesh.material_id)
        geom = _geom_e(mesh.grid)
        excluded = np.asarray(pec_e_flat, dtype=bool)
        blocks = []
        for mid, mat in disp:
            idx = np.nonzero((own == mid) & ~excluded)[0]
            if idx.size == 0:
                continue
            blocks.append((mid, idx, EPS0 * geom[idx], [_Pole(a, r, dt) for a, r in mat.poles]))
        if not blocks:
            return None
        return cls(blocks, own.size, dt)

    # -- device ---------------------------------------------------------
    def bind(self, solver) -> None:
        """Upload states once the final beta_E exists (beta sampled on the host)."""
        dims, K = solver.dims, solver.nsets
        np_dt = solver._np_dtype
        dev = solver.e.device
        beta = np.asarray(solver._coeffs.beta_e)
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

        self.idx = up(layout.index_to_device(idx, *dims, "E"), torch.int64)
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

    def save(self, solver, stream) -> None:
        _lib.call("mgh_ade_save", self.code, self.nsets, solver.e.data_ptr(), solver.e.shape[1],
                  self.n, self.idx.data_ptr(), self.fprev.data_ptr(), int(stream.cuda_stream))

    def update(self, solver, stream) -> None:
        _lib.call("mgh_ade_update", self.code, self.nsets, solver.e.data_ptr(), solver.e.shape[1],
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
