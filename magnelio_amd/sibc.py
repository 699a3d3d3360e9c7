"""Surface-impedance (SIBC) walls on the H update.

Reference: ``SIBCOperator`` (src/magnelio/solver/_sibc.py:104-264) hooked
into ``FITTimeDomainSolver`` (fit_td.py:419-429, :587-588, :839-840,
:909-916), its inputs ``SIBCSpec`` (_sibc.py:65-89), the wall enumeration
``enumerate_sibc_surfaces`` / ``resolve_wall_conductors``
(src/magnelio/mesh/_surfaces.py:758-1100) and the passive Foster fit
``fit_surface_impedance`` / ``fit_wall_impedances``
(src/magnelio/materials/surface_impedance.py:49-383).

A wall restores the Leontovich voltage E_tan = Z_s(w) (n x H) of the PEC
edges it masks as a damping term on the H faces next to it,
T_f = Z_s G_f h_f with G_f = A_f / l_dual^2.  Z_s is the ladder
c0 + sum_p c_p s/(s + b_p); its instantaneous part r_inst = c0 +
sum_p c_p / (1 + b_p dt/2) folds into M_sigma_m (W = G r_inst), the rest is
one trapezoidal state per branch, u_p+ = k_p u_p- + q_p h_mid, whose history
the device adds beta_H-weighted after the H curl (csrc/mgh_sibc.hip).

Inputs are duck-typed like the reference's: a spec with ``surfaces``
(each with ``tag``, ``comp``, ``flat_idx``, ``weight``, ``inv_l_dual``) and
``fits`` ({tag: object with ``c0`` and ``branches`` ((b_p, c_p), ...)}), so
the reference's own ``SIBCSpec`` drives this backend unchanged.  This
module also builds one from a mesh for staircase PEC solids and PEC
domain faces, and fits smooth conductors.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, layout
from .mesh import MU0

FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
TAB = 4


# ---------------------------------------------------------------------------
# spec
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class SIBCSpec:
    """Wall topology and per-tag impedance fits (_sibc.py:65-89)."""

    surfaces: tuple
    fits: dict


@dataclass
class WallSurface:
    """Driving H faces of one wall tag (mesh/_surfaces.py:758-798): per row
    the component (0 Hx, 1 Hy, 2 Hz), the C-order index inside it, the wall
    area booked onto the face and 1 / l_dual of the face."""

    tag: object
    comp: np.ndarray
    flat_idx: np.ndarray
    weight: np.ndarray
    inv_l_dual: np.ndarray

    @property
    def g(self) -> np.ndarray:
        return self.weight * self.inv_l_dual ** 2

    @property
    def area(self) -> float:
        return float(self.weight.sum()) / 2.0


@dataclass(frozen=True)
class SurfaceImpedanceFit:
    """Passive ladder c0 + sum c_p s/(s + b_p) (surface_impedance.py:166-218)."""

    sigma: float
    mu: float
    f_lo: float
    f_hi: float
    c0: float
    branches: tuple
    rel_err_re: float = field(default=0.0, compare=False)

    def impedance(self, f) -> np.ndarray:
        s = 2j * np.pi * np.asarray(f, dtype=float)
        z = np.full(s.shape, complex(self.c0))
        for b_p, c_p in self.branches:
            z = z + c_p * s / (s + b_p)
        return z

    @property
    def r_instantaneous(self) -> float:
        return self.c0 + sum(c for _, c in self.branches)


def smooth_surface_impedance(f, sigma: float, mu: float = 1.0) -> np.ndarray:
    """sqrt(j w mu0 mu / sigma) (surface_impedance.py:49-66)."""
    w = 2.0 * np.pi * np.asarray(f, dtype=float)
    return np.sqrt(1j * w * MU0 * mu / sigma)


def _ladder(w, z, n_branches: int, guard: float):
    """NNLS ladder in normalised coordinates (s/w0, Z/z0), poles log-spaced
    over the band widened by ``guard``, rows weighted by 1/|Z|."""
    from scipy.optimize import nnls  # noqa: PLC0415

    w0 = float(np.sqrt(w[0] * w[-1]))
    z0 = float(np.max(np.abs(z)))
    sn = 1j * w / w0
    bn = (2.0 * np.pi / w0) * np.logspace(np.log10(w[0] / (2.0 * np.pi) / guard),
                                          np.log10(w[-1] / (2.0 * np.pi) * guard), int(n_branches))
    zn = z / z0
    wt = 1.0 / np.abs(zn)
    m = np.column_stack([sn / (sn + b) for b in bn] + [np.ones_like(sn)]) * wt[:, None]
    coef, _ = nnls(np.vstack([m.real, m.imag]), np.concatenate([(zn * wt).real, (zn * wt).imag]),
                   maxiter=200 * m.shape[1])
    return z0 * coef[-1], bn * w0, z0 * coef[:-1]


def fit_surface_impedance(sigma: float, mu: float = 1.0, f_lo: float = 1e8, f_hi: float = 1e11,
                          tol: float = 1e-3, guard: float = 10.0,
                          max_branches: int = 32) -> SurfaceImpedanceFit:
    """Smallest even ladder (from 4 branches) whose Re Z matches the smooth
    conductor's R_s within ``tol`` on a check grid (surface_impedance.py:
    219-330; smooth conductors only)."""
    if sigma <= 0.0:
        raise ValueError(f"sigma must be > 0, got {sigma!r}")
    if not f_hi > f_lo > 0.0:
        raise ValueError(f"need 0 < f_lo < f_hi, got {f_lo!r}, {f_hi!r}")
    n_fit = max(60, int(100 * np.log10(f_hi / f_lo)))
    f_band = np.logspace(np.log10(f_lo), np.log10(f_hi), n_fit)
    f_check = np.logspace(np.log10(f_lo), np.log10(f_hi), 2 * n_fit + 1)
    z_band = smooth_surface_impedance(f_band, sigma, mu)
    r_check = smooth_surface_impedance(f_check, sigma, mu).real
    err = np.inf
    for nb in range(4, max_branches + 1, 2):
        c0, b, c = _ladder(2.0 * np.pi * f_band, z_band, nb, guard)
        fit = SurfaceImpedanceFit(float(sigma), float(mu), float(f_lo), float(f_hi), float(c0),
                                  tuple((float(bp), float(cp)) for bp, cp in zip(b, c) if cp > 0.0))
        err = float(np.max(np.abs(fit.impedance(f_check).real - r_check) / r_check))
        if err <= tol:
            return SurfaceImpedanceFit(fit.sigma, fit.mu, fit.f_lo, fit.f_hi, fit.c0, fit.branches, err)
    raise ValueError(f"surface impedance: tolerance {tol:g} not reached with {max_branches} branches "
                     f"(last error {err:.2e})")


def fit_wall_impedances(resolved: dict, f_lo: float, f_hi: float, **kw) -> dict:
    """{tag: fit}, one fit per distinct (sigma, mu) (surface_impedance.py:333-383)."""
    cache, fits = {}, {}
    for tag, (sigma, mu, *_rest) in resolved.items():
        key = (float(sigma), float(mu))
        if key not in cache:
            cache[key] = fit_surface_impedance(sigma, mu, f_lo=f_lo, f_hi=f_hi, **kw)
        fits[tag] = cache[key]
    return fits


# ---------------------------------------------------------------------------
# wall enumeration: staircase PEC solids and PEC domain faces
# ---------------------------------------------------------------------------
def _node_duals(d: np.ndarray) -> np.ndarray:
    """Dual length at every node, the first/last full cell at the ends
    (the lengths M_mu divides by, _surfaces.py:138-153)."""
    out = np.empty(d.size + 1)
    out[0], out[-1] = d[0], d[-1]
    out[1:-1] = 0.5 * (d[:-1] + d[1:])
    return out


def _live_faces(ncell, pec_mask) -> list:
    """Per H component: False where all four rim edges are PEC (the face is
    Faraday-dead, _surfaces.py:405-420)."""
    nx, ny, nz = ncell
    shapes = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    ex, ey, ez = (np.asarray(pec_mask[c, :int(np.prod(s))], bool).reshape(s) for c, s in enumerate(shapes))
    return [~(ey[:, :, :-1] & ey[:, :, 1:] & ez[:, :-1, :] & ez[:, 1:, :]),
            ~(ex[:, :, :-1] & ex[:, :, 1:] & ez[:-1, :, :] & ez[1:, :, :]),
            ~(ex[:, :-1, :] & ex[:, 1:, :] & ey[:-1, :, :] & ey[1:, :, :])]


class _Booking:
    """Wall area per H face of one tag."""

    def __init__(self, ncell):
        nx, ny, nz = ncell
        self.w = [np.zeros((nx + 1, ny, nz)), np.zeros((nx, ny + 1, nz)), np.zeros((nx, ny, nz + 1))]

    def surface(self, tag, inv_dual) -> WallSurface | None:
        rows = []
        for c in range(3):
            flat = self.w[c].ravel()
            k = np.flatnonzero(flat)
            if k.size:
                along = np.unravel_index(k, self.w[c].shape)[c]
                rows.append((np.full(k.size, c, np.uint8), k, flat[k], inv_dual[c][along]))
        if not rows:
            return None
        return WallSurface(tag, *(np.concatenate(col) for col in zip(*rows)))


def _book_walls(booking_of, tags, n, off, lengths, duals, live, ncell) -> None:
    """Book one family of walls normal to axis ``n``.  ``tags``: the wall's
    tag at every (n-node plane, tangential cell), -1 where none.  A wall edge
    along tangential axis t lies on a node of the other tangential axis u,
    between the footprints of u-cells j - 1 and j; each present footprint
    books l_t dual_u onto the live H_u face of that u-node in the driven
    n-cell (n-node + off), half of it when the edge has footprints on both
    sides (_surfaces.py:817-862)."""
    for t in range(3):
        if t == n:
            continue
        u = 3 - n - t
        nu = ncell[u]
        shape = list(tags.shape)
        shape[u] = nu + 1
        lower = np.full(shape, -1, np.int64)  # footprint in u-cell j - 1 of u-node j
        upper = np.full(shape, -1, np.int64)  # footprint in u-cell j
        dst = [slice(None)] * 3
        dst[u] = slice(1, nu + 1)
        lower[tuple(dst)] = tags
        dst[u] = slice(0, nu)
        upper[tuple(dst)] = tags
        both = (lower >= 0) & (upper >= 0)
        for foot in (lower, upper):
            at = np.nonzero(foot >= 0)
            if at[0].size == 0:
                continue
            face = list(at)
            face[n] = face[n] + off
            ok = live[u][tuple(face)]
            face = [a[ok] for a in face]
            if face[0].size == 0:
                continue
            area = np.where(both[at][ok], 0.5, 1.0) * lengths[t][face[t]] * duals[u][face[u]]
            tg = foot[at][ok]
            for tag in np.unique(tg):
                sel = tg == tag
                np.add.at(booking_of(int(tag)).w[u], tuple(a[sel] for a in face), area[sel])


def enumerate_sibc_surfaces(mesh, bc_pec_faces=()) -> list:
    """Wall surfaces of the staircase PEC solids (one per material id,
    sorted) and of the given PEC domain faces (in order), as the reference
    enumerates them for meshes without conformal data (_surfaces.py:935-1030).
    Faces that host ports must not be passed."""
    grid = mesh.grid
    lengths = (np.asarray(grid.dx), np.asarray(grid.dy), np.asarray(grid.dz))
    ncell = tuple(a.size for a in lengths)
    duals = tuple(_node_duals(a) for a in lengths)
    inv_dual = tuple(1.0 / a for a in duals)
    mat_id = mesh.material_id
    lib = mesh.material_library
    is_pec = np.zeros(max(lib) + 1, bool)
    for m, mat in lib.items():
        is_pec[m] = mat.is_pec
    pec = is_pec[mat_id]
    live = _live_faces(ncell, mesh.pec_mask_edges)
    bookings: dict = {}

    def booking_of(tag):
        return bookings.setdefault(tag, _Booking(ncell))

    for n in range(3):
        below_sl = [slice(None)] * 3
        above_sl = [slice(None)] * 3
        below_sl[n], above_sl[n] = slice(0, ncell[n] - 1), slice(1, ncell[n])
        below, above = pec[tuple(below_sl)], pec[tuple(above_sl)]
        # interior node planes 1 .. N-1: PEC below and air above drives the
        # cell above (off 0), air below and PEC above the cell below (off -1)
        for solid, wall, off in ((below_sl, below & ~above, 0), (above_sl, above & ~below, -1)):
            if not wall.any():
                continue
            shape = list(ncell)
            shape[n] += 1
            tags = np.full(shape, -1, np.int64)
            inner = [slice(None)] * 3
            inner[n] = slice(1, ncell[n])
            plane = tags[tuple(inner)]
            plane[wall] = mat_id[tuple(solid)][wall]
            _book_walls(booking_of, tags, n, off, lengths, duals, live, ncell)
    out = [s for s in (bookings[t].surface(t, inv_dual) for t in sorted(bookings)) if s is not None]
    for face in bc_pec_faces:
        if face not in FACES:
            raise ValueError(f"unknown boundary face {face!r}")
        n = "xyz".index(face[0])
        cell = 0 if face.endswith("min") else ncell[n] - 1
        sl = [slice(None)] * 3
        sl[n] = slice(cell, cell + 1)
        open_cells = ~pec[tuple(sl)]
        if not open_cells.any():
            continue
        shape = list(ncell)
        shape[n] += 1
        tags = np.full(shape, -1, np.int64)
        node = [slice(None)] * 3
        node[n] = slice(0, 1) if cell == 0 else slice(ncell[n], ncell[n] + 1)
        plane = tags[tuple(node)]
        plane[open_cells] = 0
        bk = _Booking(ncell)
        _book_walls(lambda _tag, bk=bk: bk, tags, n, 0 if cell == 0 else -1, lengths, duals, live, ncell)
        s = bk.surface(face, inv_dual)
        if s is not None:
            out.append(s)
    return out


def resolve_wall_conductors(mesh, surfaces, sigma: float | None = None, mu: float = 1.0) -> dict:
    """{tag: (sigma, mu)}: a lossy-metal solid (PEC with a finite sigma) its
    own values, every other wall the fallback (_surfaces.py:1033-1100)."""
    out = {}
    for s in surfaces:
        mat = mesh.material_library.get(s.tag) if isinstance(s.tag, (int, np.integer)) else None
        if mat is not None and mat.is_pec and mat.sigma[0] > 0.0:
            out[s.tag] = (float(mat.sigma[0]), float(mat.mu[0]))
        elif sigma is not None:
            out[s.tag] = (float(sigma), float(mu))
        else:
            raise ValueError(f"wall {s.tag!r} has no conductivity: not a lossy metal and no fallback "
                             "sigma given")
    return out


# ---------------------------------------------------------------------------
# device operator
# ---------------------------------------------------------------------------
def _h_offsets(dims) -> np.ndarray:
    nx, ny, nz = dims
    return np.array([0, (nx + 1) * ny * nz, (nx + 1) * ny * nz + nx * (ny + 1) * nz], np.int64)


class DeviceSIBC:
    """Wall blocks of a spec on the device (one block per tag with a nonzero
    impedance, in spec order); ``None`` from :meth:`from_spec` when no wall
    would act, which leaves the PEC path untouched."""

    def __init__(self, blocks, n_h: int, dt: float):
        self.blocks = blocks  # [(tag, H idx reference layout, g float64, r_inst, [(k, q)])]
        W = np.zeros(n_h)
        for _, idx, g, r_inst, _br in blocks:
            W[idx] += g * r_inst
        self.W = W
        self.dt = dt

    @classmethod
    def from_spec(cls, spec, dims, dt: float, frozen=None):
        blocks = []
        off = _h_offsets(dims)
        for surf in spec.surfaces:
            if surf.tag not in spec.fits:
                raise KeyError(f"no surface-impedance fit for wall {surf.tag!r}")
            fit = spec.fits[surf.tag]
            branches = [(float(b), float(c)) for b, c in fit.branches]
            if float(fit.c0) == 0.0 and not branches:
                continue  # Z = 0: sigma -> inf, exactly PEC
            idx = off[np.asarray(surf.comp, np.int64)] + np.asarray(surf.flat_idx, np.int64)
            g = np.asarray(surf.weight, float) * np.asarray(surf.inv_l_dual, float) ** 2
            if frozen is not None:
                keep = ~np.asarray(frozen, bool)[idx]
                idx, g = idx[keep], g[keep]
            if idx.size == 0:
                continue
            r_inst = float(fit.c0) + sum(c / (1.0 + b * dt / 2.0) for b, c in branches)
            kq = [((1.0 - b * dt / 2.0) / (1.0 + b * dt / 2.0), c * b * dt / (1.0 + b * dt / 2.0))
                  for b, c in branches]
            blocks.append((surf.tag, idx, g, r_inst, kq))
        if not blocks:
            return None
        nx, ny, nz = dims
        return cls(blocks, int(off[2]) + nx * ny * (nz + 1), dt)

    def bind(self, solver) -> None:
        dims, K = solver.dims, solver.nsets
        np_dt, t_dt = solver._np_dtype, solver.h.dtype
        dev = solver.h.device
        beta = np.asarray(solver._coeffs.beta_h)
        idx = np.concatenate([b[1] for b in self.blocks])
        dev_idx = layout.index_to_device(idx, *dims, "H")
        blk = np.concatenate([np.full(b[1].size, q, np.int32) for q, b in enumerate(self.blocks)])
        self.bmax = max(1, max(len(b[4]) for b in self.blocks))
        nbr = np.array([len(b[4]) for b in self.blocks], np.int32)
        tab = np.zeros((len(self.blocks), self.bmax, TAB), dtype=np_dt)
        for q, b in enumerate(self.blocks):
            for p, (k, qq) in enumerate(b[4]):
                tab[q, p, 0], tab[q, p, 1], tab[q, p, 2] = 0.5 * (1.0 + k), k, qq
        # distinct faces, each with its entries in block order
        faces, first, inv = np.unique(dev_idx, return_index=True, return_inverse=True)
        order = np.argsort(inv, kind="stable")
        ptr = np.zeros(faces.size + 1, np.int32)
        np.cumsum(np.bincount(inv, minlength=faces.size), out=ptr[1:])
        g = np.concatenate([b[2] for b in self.blocks]).astype(np_dt)

        def up(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)

        self.n = int(idx.size)
        self.nf = int(faces.size)
        self.idx = up(dev_idx, torch.int64)
        self.faces = up(faces, torch.int64)
        self.ptr = up(ptr, torch.int32)
        self.ent = up(order.astype(np.int32), torch.int32)
        self.blk = up(blk, torch.int32)
        self.nbr = up(nbr, torch.int32)
        self.tab = up(tab, t_dt)
        self.g = up(g, t_dt)
        self.beta = up(beta[idx].astype(np_dt), t_dt)
        self.hprev = torch.zeros((K, self.n), dtype=t_dt, device=dev)
        self.u = torch.zeros((K, self.bmax, self.n), dtype=t_dt, device=dev)
        self.code = 4 if t_dt == torch.float32 else 8
        self.nsets = K

    def save(self, solver, stream) -> None:
        _lib.call("mgh_ade_save", self.code, self.nsets, solver.h.data_ptr(), solver.h.shape[1],
                  self.n, self.idx.data_ptr(), self.hprev.data_ptr(), int(stream.cuda_stream))

    def update(self, solver, stream) -> None:
        _lib.call("mgh_sibc_update", self.code, self.nsets, solver.h.data_ptr(), solver.h.shape[1],
                  self.nf, self.faces.data_ptr(), self.ptr.data_ptr(), self.ent.data_ptr(), self.n,
                  self.idx.data_ptr(), self.blk.data_ptr(), self.g.data_ptr(), self.beta.data_ptr(),
                  self.hprev.data_ptr(), self.nbr.data_ptr(), self.tab.data_ptr(), self.bmax,
                  self.u.data_ptr(), int(stream.cuda_stream))

    def reference_state(self, s: int) -> dict:
        """{str(tag): {"u0": ..., ...}} of field set s, the reference's
        SIBCOperator.state_dict (_sibc.py:251-258)."""
        u = self.u[s].cpu().numpy()
        out, o = {}, 0
        for tag, idx, _g, _r, kq in self.blocks:
            out[str(tag)] = {f"u{p}": u[p, o:o + idx.size].copy() for p in range(len(kq))}
            o += idx.size
        return out

    def load_reference_state(self, s: int, state: dict) -> None:
        u = self.u[s].cpu().numpy()
        o = 0
        for tag, idx, _g, _r, kq in self.blocks:
            d = state[str(tag)]
            for p in range(len(kq)):
                u[p, o:o + idx.size] = np.asarray(d[f"u{p}"])
            o += idx.size
        self.u[s].copy_(torch.from_numpy(u))
