"""Oracle: trapezoidal ADE for pole-residue materials, eps(w) on the E edges
and mu(w) on the H faces (CPU).  Test infrastructure only.

Restates src/magnelio/solver/_dispersion.py (paths relative to the
reference repository root), the NumPy branch that the reference's array
backends run:

* pole coefficients  :170-186   k = (1 + a dt/2)/(1 - a dt/2), c = r/(1 - a dt/2);
                                real pole (Im a = 0): real state, weight 1;
                                otherwise a conjugate pair stored once, weight 2
* W                  :206-211   W[idx] += g * sum_p weight_p Re(c_p), folded as
                                0.5 dt W into alpha/beta (fit_td.py:382-403)
* blocks             :258-381   one block per dispersive material (sorted by
                                id): edges owned (clamped one-sided lookup)
                                by that material and not PEC; g = EPS0 A_dual/L
* H side (mu(w))     :45-76, :308-318   the same operator over the H faces:
                                owner = the clamped lookup of build_M_sigma_m,
                                g = MU0 A_primal / L_dual, frozen faces
                                (M_mu <= 0) excluded, W folded into
                                alpha_H / beta_H (fit_td.py:430-461); the
                                field stash at the top of the step next to
                                the E one, the completion right after the H
                                curl, before the CPML (fit_td.py:826-836,
                                :891-921)
* precision          :383-414   states, g, f_prev, beta in the field dtype;
                                k, c Python scalars (weak: they take the
                                array's precision)
* save / update      :466-533   f_prev = f[idx] at the top of the step; after
                                the curl: j = sum_p (0.5 w_p) Re((1 + k_p) J_p)
                                (first term seeds), f -= beta j;
                                gd = g (f - f_prev); J = J k; J += c gd
                                (complex products are NumPy's own: on FMA
                                hosts its SIMD loop rounds a*b as
                                fma(a.re, b.re, -a.im b.im),
                                fma(a.re, b.im, a.im b.re))

Pinned by tests/test_oracle_dispersion.py: a Drude DC pole equals the
sigma = eps0 r conductor (the reference's test_solver_dispersion.py:129-162
gate) and, on the H side, a mu-Drude DC pole equals the sigma* = mu0 r
magnetic conductor; each pole recursion realises the bilinear transform of
r s/(s - a) on a sampled sinusoid, and a stored-once conjugate pair equals
the explicit sum of its two complex poles.
"""

from __future__ import annotations

import numpy as np

from .mesh import EPS0, MU0, dual_widths


def geom_e(x, y, z):
    """A_dual / L_primal for every E edge (material_matrices.py:1214-1233)."""
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    ax, ay, az = dual_widths(dx), dual_widths(dy), dual_widths(dz)
    gx = (ay[None, :, None] * az[None, None, :] / dx[:, None, None]).ravel()
    gy = (ax[:, None, None] * az[None, None, :] / dy[None, :, None]).ravel()
    gz = (ax[:, None, None] * ay[None, :, None] / dz[None, None, :]).ravel()
    return np.concatenate([gx, gy, gz])


def geom_h(x, y, z):
    """A_primal / L_dual for every H face (material_matrices.py:1250-1269)."""
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    ax, ay, az = dual_widths(dx), dual_widths(dy), dual_widths(dz)
    gx = (dy[None, :, None] * dz[None, None, :] / ax[:, None, None]).ravel()
    gy = (dx[:, None, None] * dz[None, None, :] / ay[None, :, None]).ravel()
    gz = (dx[:, None, None] * dy[None, :, None] / az[None, None, :]).ravel()
    return np.concatenate([gx, gy, gz])


def face_owner(mat_id):
    """Owning cell material of every H face (_dispersion.py:309-316)."""
    nx, ny, nz = mat_id.shape
    ic = np.clip(np.arange(nx + 1), 0, max(nx - 1, 0))
    jc = np.clip(np.arange(ny + 1), 0, max(ny - 1, 0))
    kc = np.clip(np.arange(nz + 1), 0, max(nz - 1, 0))
    return np.concatenate([mat_id[ic].ravel(), mat_id[:, jc].ravel(), mat_id[:, :, kc].ravel()])


def edge_owner(mat_id):
    """Owning cell material of every E edge (clamped one-sided lookup)."""
    nx, ny, nz = mat_id.shape
    ic = np.clip(np.arange(nx + 1), 0, max(nx - 1, 0))
    jc = np.clip(np.arange(ny + 1), 0, max(ny - 1, 0))
    kc = np.clip(np.arange(nz + 1), 0, max(nz - 1, 0))
    return np.concatenate([mat_id[:, jc][:, :, kc].ravel(), mat_id[ic][:, :, kc].ravel(),
                           mat_id[ic][:, jc].ravel()])


class Pole:
    def __init__(self, a, r, dt, n, weight=None):
        k = (1.0 + a * dt / 2.0) / (1.0 - a * dt / 2.0)
        c = r / (1.0 - a * dt / 2.0)
        if complex(a).imag == 0.0 and weight is None:
            self.k, self.c, self.weight = complex(k).real, complex(c).real, 1.0
            self.J = np.zeros(n, np.float64)
        else:
            self.k, self.c = complex(k), complex(c)
            self.weight = 2.0 if weight is None else weight
            self.J = np.zeros(n, np.complex128)


class Block:
    def __init__(self, idx, g, poles):
        self.idx, self.g, self.poles = idx, g, poles
        self.beta = None
        self.f_prev = np.zeros(idx.size)


class Dispersion:
    """ADE over the dispersive materials ({mat_id: [(a, r), ...]}) of one side:
    "E" (excluded: the PEC edge mask) or "H" (excluded: the frozen faces,
    M_mu <= 0)."""

    def __init__(self, x, y, z, mat_id, poles_by_mid, dt, pec_e=None, pair_weight=None,
                 side="E", excluded=None):
        if side == "E":
            own, geom, const = edge_owner(mat_id), geom_e(x, y, z), EPS0
            excluded = pec_e if excluded is None else excluded
        else:
            own, geom, const = face_owner(mat_id), geom_h(x, y, z), MU0
        excluded = np.zeros(own.size, bool) if excluded is None else np.asarray(excluded, bool)
        self.blocks = []
        for mid in sorted(poles_by_mid):
            idx = np.nonzero((own == mid) & ~excluded)[0]
            if idx.size == 0 or not poles_by_mid[mid]:
                continue
            g = const * geom[idx]
            poles = [Pole(a, r, dt, idx.size, pair_weight) for a, r in poles_by_mid[mid]]
            self.blocks.append(Block(idx, g, poles))
        W = np.zeros(own.size)
        for b in self.blocks:
            W[b.idx] += b.g * sum(p.weight * (p.c.real if isinstance(p.c, complex) else p.c)
                                  for p in b.poles)
        self.W = W
        self.w_half = 0.5 * dt * W

    def bind(self, beta, dtype):
        cdt = np.complex64 if np.dtype(dtype).itemsize == 4 else np.complex128
        for b in self.blocks:
            b.g = b.g.astype(dtype)
            b.f_prev = b.f_prev.astype(dtype)
            b.beta = beta[b.idx].astype(dtype)
            for p in b.poles:
                p.J = p.J.astype(cdt if np.iscomplexobj(p.J) else dtype)

    def save_field(self, f):
        for b in self.blocks:
            b.f_prev[:] = f[b.idx]

    def update_field(self, f):
        for b in self.blocks:
            j_hist = None
            for p in b.poles:
                term = (0.5 * p.weight) * ((1.0 + p.k) * p.J).real
                j_hist = term if j_hist is None else j_hist + term
            f[b.idx] = f[b.idx] - b.beta * j_hist
        for b in self.blocks:
            gdelta = b.g * (f[b.idx] - b.f_prev)
            for p in b.poles:
                p.J *= p.k
                p.J += p.c * gdelta


def lorentz_pair(d_eps, w0, gamma):
    """Pole-residue pair (stored once) of d_eps w0^2 / (w0^2 - w^2 + j gamma w)."""
    wd = np.sqrt(w0 * w0 - 0.25 * gamma * gamma)
    a = complex(-0.5 * gamma, wd)
    r = complex(0.0, -d_eps * w0 * w0 / (2.0 * wd))
    return a, r
