"""Oracle: running DFT monitors (CPU, complex128).

Restates src/magnelio/monitors/_dft.py:37-52 (bins += exp(+j w t) dt data),
src/magnelio/monitors/base.py:51-173 (cell-centre physical fields: E as the
0.25-weighted sum of four edge voltages over the primal length, H as the
mean of the two face values over their solver dual widths) and the record
hooks of field_frequency.py:249-294 (E at t = n dt, H at t + dt/2) and
far_field.py:123-248 (the Huygens box placement and the tangential
components of each box face linearly weighted between the two cell layers
around the box plane).

Pinned by tests/test_oracle_monitors.py: DFT of DC and of a sampled sine
against closed forms (the reference's test_monitors.py:762-787 pair), and
the cell-centre average of a linear field reproducing the field at the
cell centre exactly; by tests/test_golden_box.py: the region and far-field
bins of a CPML / TF-SF run equal the reference's own accumulators.
"""

from __future__ import annotations

import numpy as np

from . import yee


def dual(d):
    n = d.size
    out = np.empty(n + 1)
    out[0] = d[0]
    out[n] = d[-1]
    if n > 1:
        out[1:n] = 0.5 * (d[:-1] + d[1:])
    return out


def cell_centred(e, h, x, y, z, comps, sl):
    """{comp: array over the cell slice} of physical cell-centre fields."""
    nx, ny, nz = x.size - 1, y.size - 1, z.size - 1
    Ex, Ey, Ez = yee.split(e, nx, ny, nz, "E")
    Hx, Hy, Hz = yee.split(h, nx, ny, nz, "H")
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    dxa, dya, dza = dual(dx), dual(dy), dual(dz)
    (i0, i1), (j0, j1), (k0, k1) = [(s.start, s.stop) for s in sl]
    out = {}
    for c in comps:
        if c == "Ex":
            a = Ex
            out[c] = 0.25 * (a[i0:i1, j0:j1, k0:k1] + a[i0:i1, j0 + 1:j1 + 1, k0:k1]
                             + a[i0:i1, j0:j1, k0 + 1:k1 + 1] + a[i0:i1, j0 + 1:j1 + 1, k0 + 1:k1 + 1]) / dx[i0:i1, None, None]
        elif c == "Ey":
            a = Ey
            out[c] = 0.25 * (a[i0:i1, j0:j1, k0:k1] + a[i0 + 1:i1 + 1, j0:j1, k0:k1]
                             + a[i0:i1, j0:j1, k0 + 1:k1 + 1] + a[i0 + 1:i1 + 1, j0:j1, k0 + 1:k1 + 1]) / dy[None, j0:j1, None]
        elif c == "Ez":
            a = Ez
            out[c] = 0.25 * (a[i0:i1, j0:j1, k0:k1] + a[i0 + 1:i1 + 1, j0:j1, k0:k1]
                             + a[i0:i1, j0 + 1:j1 + 1, k0:k1] + a[i0 + 1:i1 + 1, j0 + 1:j1 + 1, k0:k1]) / dz[None, None, k0:k1]
        elif c == "Hx":
            out[c] = 0.5 * (Hx[i0:i1, j0:j1, k0:k1] / dxa[i0:i1, None, None]
                            + Hx[i0 + 1:i1 + 1, j0:j1, k0:k1] / dxa[i0 + 1:i1 + 1, None, None])
        elif c == "Hy":
            out[c] = 0.5 * (Hy[i0:i1, j0:j1, k0:k1] / dya[None, j0:j1, None]
                            + Hy[i0:i1, j0 + 1:j1 + 1, k0:k1] / dya[None, j0 + 1:j1 + 1, None])
        elif c == "Hz":
            out[c] = 0.5 * (Hz[i0:i1, j0:j1, k0:k1] / dza[None, None, k0:k1]
                            + Hz[i0:i1, j0:j1, k0 + 1:k1 + 1] / dza[None, None, k0 + 1:k1 + 1])
    return out


class FieldDFT:
    """Region DFT monitor; ``record`` follows the solver's monitor hook."""

    def __init__(self, freqs, x, y, z, sl, comps=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        self.freqs = np.asarray(freqs, float)
        self.w = 2 * np.pi * self.freqs
        self.x, self.y, self.z = x, y, z
        self.sl = sl
        self.comps = list(comps)
        shape = tuple(s.stop - s.start for s in sl)
        self.bins = {c: np.zeros((self.freqs.size, *shape), complex) for c in self.comps}

    def record(self, e, h, n, t, dt):
        vals = cell_centred(e, h, self.x, self.y, self.z, self.comps, self.sl)
        for c, v in vals.items():
            tc = t + 0.5 * dt if c.startswith("H") else t
            ph = np.exp(1j * self.w * tc) * dt
            self.bins[c] += ph.reshape(-1, 1, 1, 1) * v


TANGENTIALS = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}
FACE_NAMES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


class FarFieldDFT:
    """Huygens-box DFT (far_field.py:123-248): the box sits ``depth`` =
    absorber cells + margin cells inside every open face; per open face and
    tangential component the sample is (1 - w) v[nn - 1] + w v[nn] of the
    cell-centred values of the two cell layers around the box node plane nn,
    w = (x[nn] - c[nn - 1]) / (c[nn] - c[nn - 1]) with c the cell centres,
    over the box's tangential cell range.  ``bins[face][comp]``: (F, n1, n2)."""

    def __init__(self, freqs, x, y, z, open_faces, depth):
        self.freqs = np.asarray(freqs, float)
        self.w = 2 * np.pi * self.freqs
        self.x, self.y, self.z = x, y, z
        nodes = (x, y, z)
        n_cells = tuple(a.size - 1 for a in nodes)
        lo, hi = [0, 0, 0], list(n_cells)
        for f in open_faces:
            a = "xyz".index(f[0])
            if f.endswith("min"):
                lo[a] = depth[f]
            else:
                hi[a] = n_cells[a] - depth[f]
        self.faces = []
        self.bins = {}
        for f in FACE_NAMES:
            if f not in open_faces:
                continue
            a = "xyz".index(f[0])
            nn = lo[a] if f.endswith("min") else hi[a]
            t1, t2 = [b for b in range(3) if b != a]
            cc = 0.5 * (nodes[a][:-1] + nodes[a][1:])
            wgt = float((nodes[a][nn] - cc[nn - 1]) / (cc[nn] - cc[nn - 1]))
            sl = [None, None, None]
            sl[a] = slice(nn - 1, nn + 1)
            sl[t1] = slice(lo[t1], hi[t1])
            sl[t2] = slice(lo[t2], hi[t2])
            self.faces.append((f, a, tuple(sl), wgt))
            shape = (hi[t1] - lo[t1], hi[t2] - lo[t2])
            self.bins[f] = {c: np.zeros((self.freqs.size, *shape), complex) for c in TANGENTIALS[a]}

    def record(self, e, h, n, t, dt):
        for f, a, sl, wgt in self.faces:
            vals = cell_centred(e, h, self.x, self.y, self.z, TANGENTIALS[a], sl)
            for c, v in vals.items():
                arr = np.moveaxis(v, a, 0)
                s = (1.0 - wgt) * arr[0] + wgt * arr[1]
                tc = t + 0.5 * dt if c.startswith("H") else t
                ph = np.exp(1j * self.w * tc) * dt
                self.bins[f][c] += ph.reshape(-1, 1, 1) * s


def dft_direct(values, freqs, dt, t0=0.0):
    """sum_n v[n] exp(+j w (t0 + n dt)) dt (the accumulator's convention)."""
    n = np.arange(len(values))
    w = 2 * np.pi * np.asarray(freqs, float)
    return (np.exp(1j * w[:, None] * (t0 + n[None, :] * dt)) * dt) @ np.asarray(values, float)
