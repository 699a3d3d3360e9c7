"""Oracle: TF/SF plane-wave injection (CPU).  Test infrastructure only.

Restates src/magnelio/sources/plane_wave.py (paths relative to the
reference repository root):

* waveform            :185-198  amplitude * exp(-(t-t0)^2 / (2 sigma^2)),
                                t0 = 4/f_max, sigma = 2/(pi f_max); or
                                amplitude * sin(2 pi f t).  The amplitude is
                                also folded into the incident vectors
                                (:249-252), so it enters squared — kept.
* box snapping        :262-306
* face corrections    :318-707  twelve E and twelve H face patches,
                                field[patch] += beta * (factor * metric) * f(t - k.r / c0)
* injection times     :710-734  E at t_E - dt/2, H at t_H - dt/2

The patch list below is written out face by face (z, x, y; min then max)
with the reference's component order and signs.  Pinned by
tests/test_oracle_tfsf.py: in an empty PEC box the incident pulse reaches
unit amplitude inside the total-field region and the scattered-field
region stays at the dispersion floor on all six propagation axes (the
reference's own gate, test_plane_wave.py:166-172: |inside - 1| < 1e-3,
leakage < 1e-4) — a sign, index, metric or delay slip on any face breaks
one of the two.
"""

from __future__ import annotations

import math

import numpy as np

from . import yee
from .mesh import C0, MU0, dual_widths

ETA0 = MU0 * C0


class PlaneWave:
    def __init__(self, x, y, z, beta_e, beta_h, dt, *, direction=(0, 0, 1), polarization=(1, 0, 0),
                 box=None, amplitude=1.0, waveform="gaussian", f_max=None, f_center=None,
                 dtype=np.float64):
        self.x, self.y, self.z = (np.asarray(a, float) for a in (x, y, z))
        self.nx, self.ny, self.nz = self.x.size - 1, self.y.size - 1, self.z.size - 1
        self.dt = dt
        self.dtype = dtype
        k = np.asarray(direction, float)
        k = k / np.linalg.norm(k)
        p = np.asarray(polarization, float)
        p = p - np.dot(p, k) * k
        p = p / np.linalg.norm(p)
        self.k = k
        self.amplitude = amplitude
        self.waveform = waveform
        self.f_max, self.f_center = f_max, f_center
        self.eh = p * amplitude
        self.hh = np.cross(k, p) * (amplitude / ETA0)
        self.box = box if box is not None else (2, self.nx - 1, 2, self.ny - 1, 2, self.nz - 1)
        self.be = beta_e.astype(dtype)
        self.bh = beta_h.astype(dtype)
        self.E, self.H = self._patches()

    def excitation(self, t):
        if self.waveform == "gaussian":
            f_max = self.f_max or 1e9
            t0 = 4.0 / f_max
            sigma = 2.0 / (math.pi * f_max)
            return self.amplitude * np.exp(-((t - t0) ** 2) / (2 * sigma**2))
        f = self.f_center or 1e9
        return self.amplitude * np.sin(2 * math.pi * f * t)

    def _patches(self):
        x, y, z = self.x, self.y, self.z
        nx, ny, nz = self.nx, self.ny, self.nz
        i0, i1, j0, j1, k0, k1 = self.box
        xc, yc, zc = ((a[:-1] + a[1:]) / 2 for a in (x, y, z))
        dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
        dxa, dya, dza = dual_widths(dx), dual_widths(dy), dual_widths(dz)
        be = dict(zip(("Ex", "Ey", "Ez"), yee.split(self.be, nx, ny, nz, "E")))
        bh = dict(zip(("Hx", "Hy", "Hz"), yee.split(self.bh, nx, ny, nz, "H")))
        ic, inn = slice(i0, i1), slice(i0, i1 + 1)
        jc, jn = slice(j0, j1), slice(j0, j1 + 1)
        kc, kn = slice(k0, k1), slice(k0, k1 + 1)
        xlo, ylo, zlo = (x[i0 - 1] + x[i0]) / 2, (y[j0 - 1] + y[j0]) / 2, (z[k0 - 1] + z[k0]) / 2
        xhi = (x[i1] + x[i1 + 1]) / 2 if i1 < nx else x[i1]
        yhi = (y[j1] + y[j1 + 1]) / 2 if j1 < ny else y[j1]
        zhi = (z[k1] + z[k1 + 1]) / 2 if k1 < nz else z[k1]
        hh, eh = self.hh, self.eh
        col, row = (lambda a: a[:, None]), (lambda a: a[None, :])
        # (component, index, metric, factor, (x, y, z) positions)
        E = [
            ("Ex", (ic, jn, k0), row(dya[jn]), +hh[1], (col(xc[ic]), row(y[jn]), zlo)),
            ("Ey", (inn, jc, k0), col(dxa[inn]), -hh[0], (col(x[inn]), row(yc[jc]), zlo)),
            ("Ex", (ic, jn, k1), row(dya[jn]), -hh[1], (col(xc[ic]), row(y[jn]), zhi)),
            ("Ey", (inn, jc, k1), col(dxa[inn]), +hh[0], (col(x[inn]), row(yc[jc]), zhi)),
            ("Ey", (i0, jc, kn), row(dza[kn]), +hh[2], (xlo, col(yc[jc]), row(z[kn]))),
            ("Ez", (i0, jn, kc), col(dya[jn]), -hh[1], (xlo, col(y[jn]), row(zc[kc]))),
            ("Ey", (i1, jc, kn), row(dza[kn]), -hh[2], (xhi, col(yc[jc]), row(z[kn]))),
            ("Ez", (i1, jn, kc), col(dya[jn]), +hh[1], (xhi, col(y[jn]), row(zc[kc]))),
            ("Ex", (ic, j0, kn), row(dza[kn]), -hh[2], (col(xc[ic]), ylo, row(z[kn]))),
            ("Ez", (inn, j0, kc), col(dxa[inn]), +hh[0], (col(x[inn]), ylo, row(zc[kc]))),
            ("Ex", (ic, j1, kn), row(dza[kn]), +hh[2], (col(xc[ic]), yhi, row(z[kn]))),
            ("Ez", (inn, j1, kc), col(dxa[inn]), -hh[0], (col(x[inn]), yhi, row(zc[kc]))),
        ]
        H = [("Hy", (ic, jn, k0 - 1), col(dx[ic]), +eh[0], (col(xc[ic]), row(y[jn]), z[k0])),
             ("Hx", (inn, jc, k0 - 1), row(dy[jc]), -eh[1], (col(x[inn]), row(yc[jc]), z[k0]))]
        if k1 < nz:
            H += [("Hy", (ic, jn, k1), col(dx[ic]), -eh[0], (col(xc[ic]), row(y[jn]), z[k1])),
                  ("Hx", (inn, jc, k1), row(dy[jc]), +eh[1], (col(x[inn]), row(yc[jc]), z[k1]))]
        H += [("Hz", (i0 - 1, jc, kn), col(dy[jc]), +eh[1], (x[i0], col(yc[jc]), row(z[kn]))),
              ("Hy", (i0 - 1, jn, kc), row(dz[kc]), -eh[2], (x[i0], col(y[jn]), row(zc[kc])))]
        if i1 < nx:
            H += [("Hz", (i1, jc, kn), col(dy[jc]), -eh[1], (x[i1], col(yc[jc]), row(z[kn]))),
                  ("Hy", (i1, jn, kc), row(dz[kc]), +eh[2], (x[i1], col(y[jn]), row(zc[kc])))]
        H += [("Hz", (ic, j0 - 1, kn), col(dx[ic]), -eh[0], (col(xc[ic]), y[j0], row(z[kn]))),
              ("Hx", (inn, j0 - 1, kc), row(dz[kc]), +eh[2], (col(x[inn]), y[j0], row(zc[kc])))]
        if j1 < ny:
            H += [("Hz", (ic, j1, kn), col(dx[ic]), +eh[0], (col(xc[ic]), y[j1], row(z[kn]))),
                  ("Hx", (inn, j1, kc), row(dz[kc]), -eh[2], (col(x[inn]), y[j1], row(zc[kc])))]

        def fold(rows, beta):
            out = []
            for comp, index, metric, factor, pos in rows:
                if factor == 0.0:
                    continue
                coef = beta[comp][index] * np.asarray(factor * metric, dtype=self.dtype)
                p = 0.0
                for kc_, c in zip(self.k, pos):
                    if kc_ != 0.0:
                        p = p + kc_ * c
                delay = np.asarray(p, dtype=np.float64) / C0
                out.append((comp, index, delay if delay.ndim else float(delay), coef))
            return out

        return fold(E, be), fold(H, bh)

    def _apply(self, patches, views, t):
        sc = np.dtype(self.dtype).type
        for comp, index, delay, coef in patches:
            w = self.excitation(t - delay)
            w = w.astype(self.dtype, copy=False) if getattr(w, "ndim", 0) else sc(w)
            views[comp][index] += coef * w

    def inject_e(self, e, t_e):
        views = dict(zip(("Ex", "Ey", "Ez"), yee.split(e, self.nx, self.ny, self.nz, "E")))
        self._apply(self.E, views, t_e - self.dt / 2)

    def inject_h(self, h, t_h):
        views = dict(zip(("Hx", "Hy", "Hz"), yee.split(h, self.nx, self.ny, self.nz, "H")))
        self._apply(self.H, views, t_h - self.dt / 2)


def snap_box(x, y, z, corners):
    """Node-index TF box from two corners (plane_wave.py:262-306)."""
    n = (x.size - 1, y.size - 1, z.size - 1)
    if corners is None:
        return (2, n[0] - 1, 2, n[1] - 1, 2, n[2] - 1)
    out = []
    for ax, nodes in enumerate((x, y, z)):
        a, b = corners[0][ax], corners[1][ax]
        a = None if a is None or not np.isfinite(a) else float(a)
        b = None if b is None or not np.isfinite(b) else float(b)
        if a is not None and b is not None and b < a:
            a, b = b, a
        lo = 2 if a is None else int(np.searchsorted(nodes, a).clip(1, n[ax] - 1))
        hi = n[ax] - 1 if b is None else int(np.searchsorted(nodes, b).clip(lo + 1, n[ax]))
        out += [lo, hi]
    return tuple(out)
