"""Oracle: CFS-PML profiles and psi recursion (CPU, NumPy).

Restates src/magnelio/boundaries/cpml.py:
  profiles      initialize  :92-195  sigma = sigma_max rho^m,
                kappa = 1 + (kappa_max-1) rho^m, alpha = alpha_max (1-rho),
                sampled at cell centres; sigma_max = -(m+1) C0 EPS0 ln(R)/(2 d)
                b = exp(-(sigma/kappa + alpha) dt/EPS0),
                c = sigma/(kappa sigma + kappa^2 alpha) (1-b) (0 where sigma=0),
                ck = 1 - 1/kappa
  air masks     set_pec_mask :213-310 (H masks: a face is masked when either
                cell on its normal is PEC)
  E correction  update_E :343-438, H correction update_H :440-516

Operation order per slab element (field dtype):
  psi = psi * b;  psi = psi + (c * d);  psi = psi * air
  F   = F +/- (beta * ((ck * d) + psi))
The outer-face PEC ``apply`` (:316-341) is NOT part of the step: the solver
never calls it (src/magnelio/solver/fit_td.py:870-882).

Parity pins (tests/test_oracle_cpml.py): closed-form profile values, the
loop-form brute force of one slab update on a small grid (bit-exact), and a
normal-incidence reflection check of a plane pulse against R_target.
"""

from __future__ import annotations

import numpy as np

from . import yee

C0 = 299_792_458.0
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0 * C0)

# face -> (normal axis, side)
FACE_AXIS = {"xmin": (0, "min"), "xmax": (0, "max"), "ymin": (1, "min"),
             "ymax": (1, "max"), "zmin": (2, "min"), "zmax": (2, "max")}

# For each normal axis: (E comp 1, H comp used by it, sign), (E comp 2, H comp, sign)
# and for H: (H comp 1, E comp used, sign), (H comp 2, E comp, sign)
#   z: Ex += (dHy), Ey -= (dHx);  Hx -= (dEy), Hy += (dEx)
#   x: Ey += (dHz), Ez -= (dHy);  Hy -= (dEz), Hz += (dEy)
#   y: Ex -= (dHz), Ez += (dHx);  Hx += (dEz), Hz -= (dEx)
E_TERMS = {2: ((0, 1, +1), (1, 0, -1)), 0: ((1, 2, +1), (2, 1, -1)), 1: ((0, 2, -1), (2, 0, +1))}
H_TERMS = {2: ((0, 1, -1), (1, 0, +1)), 0: ((1, 2, -1), (2, 1, +1)), 1: ((0, 2, +1), (2, 0, -1))}


def profiles(d_axis: np.ndarray, side: str, dt: float, n_layers: int = 8, m: int = 3,
             r_target: float = 1e-8, kappa_max: float = 7.0, alpha_max: float = 0.02):
    """(b, c, ck, cells) per PML layer, ordered by increasing global index."""
    n = d_axis.size
    npml = min(n_layers, n)
    cells = list(range(npml)) if side == "min" else list(range(n - npml, n))
    d_phys = 0.0
    for k in cells:  # in-order scalar sum, as the reference's builtin sum
        d_phys = d_phys + float(d_axis[k])
    d_phys = d_phys or 1.0
    sigma_max = -(m + 1) * C0 * EPS0 * np.log(r_target) / (2 * d_phys)
    sigma, kappa, alpha = np.zeros(npml), np.ones(npml), np.zeros(npml)
    cum = 0.0
    for li, ci in enumerate(cells):
        half = d_axis[ci] / 2.0
        rho = 1.0 - (cum + half) / d_phys if side == "min" else (cum + half) / d_phys
        cum += d_axis[ci]
        sigma[li] = sigma_max * rho**m
        kappa[li] = 1.0 + (kappa_max - 1.0) * rho**m
        alpha[li] = alpha_max * (1.0 - rho)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS0)
    c = np.where(sigma > 0.0, sigma / (kappa * sigma + kappa**2 * alpha) * (1.0 - b), 0.0)
    ck = 1.0 - 1.0 / kappa
    return b, c, ck, cells


def air_masks(pec_e_flat: np.ndarray, pec_cell: np.ndarray, nx, ny, nz, axis, cells):
    """0/1 masks (E1, E2, H1, H2) over the slab, shaped like the psi arrays."""
    ex, ey, ez = yee.split(pec_e_flat, nx, ny, nz, "E")
    phx = np.zeros((nx + 1, ny, nz), bool)
    phy = np.zeros((nx, ny + 1, nz), bool)
    phz = np.zeros((nx, ny, nz + 1), bool)
    if pec_cell.any():
        phx[:-1] |= pec_cell
        phx[1:] |= pec_cell
        phy[:, :-1] |= pec_cell
        phy[:, 1:] |= pec_cell
        phz[:, :, :-1] |= pec_cell
        phz[:, :, 1:] |= pec_cell
    e_comp = [ex, ey, ez]
    h_comp = [phx, phy, phz]
    (e1, _, _), (e2, _, _) = E_TERMS[axis]
    (h1, _, _), (h2, _, _) = H_TERMS[axis]

    def take(a):
        return (~np.take(a, cells, axis=axis)).astype(np.float64)

    return take(e_comp[e1]), take(e_comp[e2]), take(h_comp[h1]), take(h_comp[h2])


class CPMLFace:
    """One CPML face: profiles, psi state and the per-step corrections."""

    def __init__(self, face, x, y, z, dt, dtype, pec_e_flat=None, pec_cell=None, **kw):
        self.face = face
        self.axis, self.side = FACE_AXIS[face]
        self.dims = (x.size - 1, y.size - 1, z.size - 1)
        d_axis = np.diff((x, y, z)[self.axis])
        b, c, ck, cells = profiles(d_axis, self.side, dt, **kw)
        self.cells = cells
        self.n = len(cells)
        self.k0 = cells[0]
        shp = [1, 1, 1]
        shp[self.axis] = self.n
        self.b = b.astype(dtype).reshape(shp)
        self.c = c.astype(dtype).reshape(shp)
        self.ck = ck.astype(dtype).reshape(shp)
        nx, ny, nz = self.dims
        esh = yee.shapes(nx, ny, nz)
        names_e = ("Ex", "Ey", "Ez")
        names_h = ("Hx", "Hy", "Hz")

        def slab_shape(name):
            s = list(esh[name])
            s[self.axis] = self.n
            return tuple(s)

        (e1, _, _), (e2, _, _) = E_TERMS[self.axis]
        (h1, _, _), (h2, _, _) = H_TERMS[self.axis]
        self.psi_e = [np.zeros(slab_shape(names_e[e1]), dtype), np.zeros(slab_shape(names_e[e2]), dtype)]
        self.psi_h = [np.zeros(slab_shape(names_h[h1]), dtype), np.zeros(slab_shape(names_h[h2]), dtype)]
        if pec_e_flat is not None:
            pc = pec_cell if pec_cell is not None else np.zeros(self.dims, bool)
            self.air = air_masks(pec_e_flat, pc, nx, ny, nz, self.axis, cells)
        else:
            self.air = None

    def _sl(self, lo, hi):
        s = [slice(None)] * 3
        s[self.axis] = slice(lo, hi)
        return tuple(s)

    def _back_diff(self, H):
        """Backward difference of H over the slab (H[-1] = 0 on the min side)."""
        n, k0 = self.n, self.k0
        if self.side == "min":
            d = H[self._sl(0, n)].copy()
            d[self._sl(1, n)] -= H[self._sl(0, n - 1)]
            return d
        return H[self._sl(k0, k0 + n)] - H[self._sl(k0 - 1, k0 + n - 1)]

    def update_e(self, e, h, beta_e):
        nx, ny, nz = self.dims
        E = yee.split(e, nx, ny, nz, "E")
        H = yee.split(h, nx, ny, nz, "H")
        B = yee.split(beta_e, nx, ny, nz, "E")
        for t, (ec, hc, sign) in enumerate(E_TERMS[self.axis]):
            d = self._back_diff(H[hc])
            psi = self.psi_e[t]
            psi *= self.b
            psi += self.c * d
            if self.air is not None:
                psi *= self.air[t]
            sl = self._sl(self.k0, self.k0 + self.n)
            corr = B[ec][sl] * (self.ck * d + psi)
            if sign > 0:
                E[ec][sl] += corr
            else:
                E[ec][sl] -= corr

    def update_h(self, e, h, beta_h):
        nx, ny, nz = self.dims
        E = yee.split(e, nx, ny, nz, "E")
        H = yee.split(h, nx, ny, nz, "H")
        B = yee.split(beta_h, nx, ny, nz, "H")
        k0, n = self.k0, self.n
        for t, (hc, ec, sign) in enumerate(H_TERMS[self.axis]):
            d = E[ec][self._sl(k0 + 1, k0 + n + 1)] - E[ec][self._sl(k0, k0 + n)]
            psi = self.psi_h[t]
            psi *= self.b
            psi += self.c * d
            if self.air is not None:
                psi *= self.air[2 + t]
            sl = self._sl(k0, k0 + n)
            corr = B[hc][sl] * (self.ck * d + psi)
            if sign > 0:
                H[hc][sl] += corr
            else:
                H[hc][sl] -= corr
