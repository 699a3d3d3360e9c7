"""Oracle: time-domain surface-impedance walls (SIBC) on the H update (CPU).
Test infrastructure only.

Restates, paths relative to the reference repository root:

* smooth impedance   materials/surface_impedance.py:49-66   Z = sqrt(j w mu0 mu / sigma)
* Foster NNLS fit    :139-163, :219-330   one ladder c0 + sum c_p s/(s + b_p) per
                     branch count 4, 6, ... until max |Re Z - R_s| / R_s <= tol on a
                     check grid twice as dense; poles log-spaced over the band
                     widened by `guard`, coefficients by NNLS in normalised
                     coordinates (w -> w/w0, Z -> Z/z0), weighted by 1/|Z|
* wall enumeration   mesh/_surfaces.py:138-153 (dual lengths, first/last full
                     cell), :405-420 (a face is alive unless its four rim edges
                     are all PEC), :817-862 (wall-edge -> driving-face pairs),
                     :865-905 (staircase PEC solids, one tag per material id),
                     :908-932 (PEC domain-boundary walls, tag = face name),
                     :758-798 (G_f = weight / l_dual^2, state indices [Hx|Hy|Hz])
* operator           solver/_sibc.py:104-249   W = G r_inst folded into M_sigma_m;
                     save h at the top of the step; after the H curl, block by
                     block h += beta (g sum_p ((1 + k_p)/2) u_p) (the first term
                     seeds), then for every block h_mid = 0.5 (h + h_prev),
                     u_p = u_p k_p, u_p += q_p h_mid; k = (1 - b dt/2)/(1 + b dt/2),
                     q = c b dt/(1 + b dt/2), r_inst = c0 + sum c_p/(1 + b_p dt/2);
                     g, u, h_prev, beta in the field dtype, k, q and the hist
                     weights Python floats (they take the array's precision)

Pinned by tests/test_oracle_sibc.py: the operator against a per-step
numerical solve of the implicit face/branch system written from the
impedance itself (one face, two branches; a seam face booked by two walls);
Z = 0 leaves no operator (the PEC path untouched); the enumeration against
the reference's own surfaces and fits (tests/golden/sibc_cavity.npz) and
the whole march against the reference's run of the same cavity.
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import nnls

MU0 = 1.25663706212e-6
FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


# ---------------------------------------------------------------------------
# impedance and its passive rational fit
# ---------------------------------------------------------------------------
def smooth_surface_impedance(f, sigma: float, mu: float = 1.0):
    w = 2.0 * np.pi * np.asarray(f, dtype=float)
    return np.sqrt(1j * w * MU0 * mu / sigma)


class Fit:
    """c0 + sum c_p s/(s + b_p) with branches ((b_p, c_p), ...)."""

    def __init__(self, c0: float, branches=()):
        self.c0 = float(c0)
        self.branches = tuple((float(b), float(c)) for b, c in branches)

    def impedance(self, f):
        s = 2j * np.pi * np.asarray(f, dtype=float)
        z = np.full(s.shape, complex(self.c0))
        for b, c in self.branches:
            z = z + c * s / (s + b)
        return z


def foster_ladder(w_band, z_target, n_branches: int, guard: float):
    """(c0, b, c) of one NNLS ladder, physical scales."""
    w0 = float(np.sqrt(w_band[0] * w_band[-1]))
    z0 = float(np.max(np.abs(z_target)))
    s = 1j * w_band / w0
    f_poles = np.logspace(np.log10(w_band[0] / (2 * np.pi) / guard),
                          np.log10(w_band[-1] / (2 * np.pi) * guard), int(n_branches))
    b = f_poles * (2 * np.pi / w0)
    zn = z_target / z0
    wt = 1.0 / np.abs(zn)
    cols = np.column_stack([s / (s + bp) for bp in b] + [np.ones_like(s)]) * wt[:, None]
    a = np.vstack([cols.real, cols.imag])
    rhs = np.concatenate([(zn * wt).real, (zn * wt).imag])
    coef, _ = nnls(a, rhs, maxiter=200 * a.shape[1])
    return z0 * coef[-1], b * w0, z0 * coef[:-1]


def fit_surface_impedance(sigma: float, mu: float = 1.0, f_lo: float = 1e8, f_hi: float = 1e11,
                          tol: float = 1e-3, guard: float = 10.0, max_branches: int = 32) -> Fit:
    decades = np.log10(f_hi / f_lo)
    n_fit = max(60, int(100 * decades))
    f_band = np.logspace(np.log10(f_lo), np.log10(f_hi), n_fit)
    f_chk = np.logspace(np.log10(f_lo), np.log10(f_hi), 2 * n_fit + 1)
    z_band = smooth_surface_impedance(f_band, sigma, mu)
    r_chk = smooth_surface_impedance(f_chk, sigma, mu).real
    for nb in range(4, max_branches + 1, 2):
        c0, b, c = foster_ladder(2 * np.pi * f_band, z_band, nb, guard)
        fit = Fit(c0, [(bp, cp) for bp, cp in zip(b, c) if cp > 0.0])
        if np.max(np.abs(fit.impedance(f_chk).real - r_chk) / r_chk) <= tol:
            return fit
    raise ValueError("surface impedance fit: tolerance not reached")


# ---------------------------------------------------------------------------
# wall enumeration (staircase PEC solids and PEC domain-boundary faces)
# ---------------------------------------------------------------------------
class Surface:
    def __init__(self, tag, comp, flat_idx, weight, inv_l_dual):
        self.tag = tag
        self.comp = np.asarray(comp, np.uint8)
        self.flat_idx = np.asarray(flat_idx, np.int64)
        self.weight = np.asarray(weight, float)
        self.inv_l_dual = np.asarray(inv_l_dual, float)

    @property
    def g(self):
        return self.weight * self.inv_l_dual ** 2

    def state_indices(self, dims):
        nx, ny, nz = dims
        off = np.array([0, (nx + 1) * ny * nz, (nx + 1) * ny * nz + nx * (ny + 1) * nz], np.int64)
        return off[self.comp] + self.flat_idx


def state_dual_lengths(d):
    d = np.asarray(d, float)
    out = np.empty(d.size + 1)
    out[0], out[-1] = d[0], d[-1]
    out[1:-1] = 0.5 * (d[:-1] + d[1:])
    return out


def h_shapes(dims):
    nx, ny, nz = dims
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]


def face_alive(dims, pec_edges):
    """Per H component: the face has at least one non-PEC rim edge.
    ``pec_edges``: flat bool [Ex|Ey|Ez] (C order per component)."""
    nx, ny, nz = dims
    sh = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    parts, o = [], 0
    for s in sh:
        n = int(np.prod(s))
        parts.append(np.asarray(pec_edges[o:o + n], bool).reshape(s))
        o += n
    ex, ey, ez = parts
    hx = ~(ey[:, :, :-1] & ey[:, :, 1:] & ez[:, :-1, :] & ez[:, 1:, :])
    hy = ~(ex[:, :, :-1] & ex[:, :, 1:] & ez[:-1, :, :] & ez[1:, :, :])
    hz = ~(ex[:, :-1, :] & ex[:, 1:, :] & ey[:-1, :, :] & ey[1:, :, :])
    return [hx, hy, hz]


class _Acc:
    def __init__(self, dims):
        self.w = [np.zeros(s) for s in h_shapes(dims)]

    def surface(self, tag, inv_dual):
        comp, idx, wt, inv = [], [], [], []
        for c in range(3):
            flat = self.w[c].ravel()
            nzi = np.nonzero(flat)[0]
            if nzi.size == 0:
                continue
            comp.append(np.full(nzi.size, c))
            idx.append(nzi)
            wt.append(flat[nzi])
            inv.append(inv_dual[c][np.unravel_index(nzi, self.w[c].shape)[c]])
        if not comp:
            return None
        return Surface(tag, np.concatenate(comp), np.concatenate(idx), np.concatenate(wt),
                       np.concatenate(inv))


def _book(acc_for, tags, n_ax, off, d, dual, alive, ncell):
    """One wall family: ``tags`` over (n-node planes) x (tangential cells),
    -1 where no wall.  Every wall edge along a tangential axis t sits at a
    node of the other tangential axis u between two footprints; each
    footprint present books l_t * dual_u (half of it when both are) onto
    the H_u face at that u-node in the driven n-cell (n-node + off)."""
    for t in range(3):
        if t == n_ax:
            continue
        u = 3 - n_ax - t
        pad = [(0, 0)] * 3
        pad[u] = (1, 1)
        p = np.pad(tags, pad, constant_values=-1)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[u] = slice(0, ncell[u] + 1)   # footprint in u-cell j - 1
        hi[u] = slice(1, ncell[u] + 2)   # footprint in u-cell j
        fa, fb = p[tuple(lo)], p[tuple(hi)]
        pair = (fa >= 0) & (fb >= 0)
        for foot in (fa, fb):
            sel = foot >= 0
            if not sel.any():
                continue
            pos = list(np.nonzero(sel))
            pos[n_ax] = pos[n_ax] + off
            live = alive[u][tuple(pos)]
            if not live.any():
                continue
            pos = [q[live] for q in pos]
            w = np.where(pair[sel][live], 0.5, 1.0) * d[t][pos[t]] * dual[u][pos[u]]
            tg = foot[sel][live]
            for tag in np.unique(tg):
                m = tg == tag
                np.add.at(acc_for(int(tag)).w[u], tuple(q[m] for q in pos), w[m])


def enumerate_surfaces(x, y, z, mat_id, pec_table, pec_edges, bc_faces=()):
    """Staircase PEC solids (one surface per material id, sorted) then the
    requested PEC domain-boundary faces (in the given order)."""
    d = (np.diff(x), np.diff(y), np.diff(z))
    ncell = tuple(a.size for a in d)
    dual = tuple(state_dual_lengths(a) for a in d)
    inv_dual = tuple(1.0 / a for a in dual)
    pec = np.asarray(pec_table, bool)[mat_id]
    alive = face_alive(ncell, pec_edges)
    accs = {}

    def acc_for(tag):
        return accs.setdefault(tag, _Acc(ncell))

    for n in range(3):
        shp = list(ncell)
        shp[n] += 1
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        mid = [slice(None)] * 3
        a[n], b[n], mid[n] = slice(None, -1), slice(1, None), slice(1, -1)
        below, above = pec[tuple(a)], pec[tuple(b)]
        for side, wall, off in ((tuple(a), below & ~above, 0), (tuple(b), above & ~below, -1)):
            if not wall.any():
                continue
            tags = np.full(shp, -1, np.int64)
            inner = tags[tuple(mid)]
            inner[wall] = mat_id[side][wall]
            _book(acc_for, tags, n, off, d, dual, alive, ncell)
    out = [s for s in (accs[t].surface(t, inv_dual) for t in sorted(accs)) if s is not None]
    for face in bc_faces:
        n = "xyz".index(face[0])
        layer = 0 if face.endswith("min") else ncell[n] - 1
        sl = [slice(None)] * 3
        sl[n] = slice(layer, layer + 1)
        open_cells = ~pec[tuple(sl)]
        if not open_cells.any():
            continue
        shp = list(ncell)
        shp[n] += 1
        tags = np.full(shp, -1, np.int64)
        node = 0 if layer == 0 else layer + 1
        pl = [slice(None)] * 3
        pl[n] = slice(node, node + 1)
        plane = tags[tuple(pl)]
        plane[open_cells] = 0
        acc = _Acc(ncell)
        _book(lambda _t, acc=acc: acc, tags, n, 0 if layer == 0 else -1, d, dual, alive, ncell)
        s = acc.surface(face, inv_dual)
        if s is not None:
            out.append(s)
    return out


# ---------------------------------------------------------------------------
# the operator
# ---------------------------------------------------------------------------
class _Block:
    def __init__(self, tag, idx, g, fit, dt):
        self.tag, self.idx, self.g = tag, idx, g
        self.r_inst = fit.c0 + sum(c / (1.0 + b * dt / 2.0) for b, c in fit.branches)
        self.k = [(1.0 - b * dt / 2.0) / (1.0 + b * dt / 2.0) for b, _ in fit.branches]
        self.q = [c * b * dt / (1.0 + b * dt / 2.0) for b, c in fit.branches]
        self.u = [np.zeros(idx.size) for _ in fit.branches]
        self.beta = None
        self.h_prev = np.zeros(idx.size)


class SIBC:
    """Wall blocks of (surfaces, {tag: Fit}); ``frozen``: H states to leave
    out (M_mu <= 0).  ``None`` from :meth:`build` when no wall has a
    nonzero impedance."""

    def __init__(self, blocks, n_h):
        self.blocks = blocks
        W = np.zeros(n_h)
        for b in blocks:
            W[b.idx] += b.g * b.r_inst
        self.W = W

    @classmethod
    def build(cls, surfaces, fits, dims, dt, frozen=None):
        blocks = []
        for s in surfaces:
            fit = fits[s.tag]
            if fit.c0 == 0.0 and not fit.branches:
                continue
            idx = s.state_indices(dims)
            g = np.asarray(s.g, float)
            if frozen is not None:
                keep = ~np.asarray(frozen, bool)[idx]
                idx, g = idx[keep], g[keep]
            if idx.size:
                blocks.append(_Block(s.tag, idx, g, fit, dt))
        if not blocks:
            return None
        nx, ny, nz = dims
        return cls(blocks, (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1))

    def bind(self, beta, dtype):
        for b in self.blocks:
            b.g = b.g.astype(dtype)
            b.h_prev = b.h_prev.astype(dtype)
            b.beta = beta[b.idx].astype(dtype)
            b.u = [u.astype(dtype) for u in b.u]

    def save_field(self, h):
        for b in self.blocks:
            b.h_prev[:] = h[b.idx]

    def update_field(self, h):
        for b in self.blocks:
            hist = None
            for k, u in zip(b.k, b.u):
                term = (0.5 * (1.0 + k)) * u
                hist = term if hist is None else hist + term
            if hist is not None:
                h[b.idx] = h[b.idx] + b.beta * (b.g * hist)
        for b in self.blocks:
            if not b.u:
                continue
            h_mid = 0.5 * (h[b.idx] + b.h_prev)
            for i, (k, q) in enumerate(zip(b.k, b.q)):
                b.u[i] = b.u[i] * k
                b.u[i] = b.u[i] + q * h_mid

    def state_dict(self):
        return {str(b.tag): {f"u{i}": u.copy() for i, u in enumerate(b.u)} for b in self.blocks}

    def load_state_dict(self, sd):
        for b in self.blocks:
            for i in range(len(b.u)):
                b.u[i][:] = sd[str(b.tag)][f"u{i}"]
