"""Host-side mesh and material-coefficient setup (float64, NumPy).

These are the inputs the stepping engine consumes; they run once per job
and never on the device hot path.  All flat arrays use the reference's
layout (component blocks, z fastest) and are moved to the device layout by
:mod:`magnelio_amd.layout` at upload.

Reference counterparts (paths relative to the reference repository root):

* ``GridLines``            src/magnelio/mesh/grid.py:17-120
* ``Mesh.from_grid``       src/magnelio/mesh/mesher.py:199-321 (AABB fill by
                           cell centre, later regions overwrite earlier ones)
* ``build_pec_mask``       src/magnelio/mesh/indexing.py:168-257 (an edge is
                           PEC when ANY adjacent cell is PEC — the code ORs)
* ``or_bbox_pec_walls``    src/magnelio/mesh/mesher.py:1370-1430
* ``build_m_eps`` etc.     src/magnelio/_operators/material_matrices.py:55-73,
                           :334-464, :494-613, :616-686, :688-757 (category-0
                           staircase: one clamped owning cell per edge/face)
* ``flatten_port_*``       material_matrices.py:106-179, :181-255, :257-331
* ``update_coefficients``  src/magnelio/solver/fit_td.py:375-513
* ``courant_dt``           src/magnelio/solver/stability.py:44-93
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C0 = 299_792_458.0
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0 = MU0 * C0

SAFETY_FACTORS = {"draft": 0.90, "normal": 0.95, "high": 0.99}
FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


@dataclass(frozen=True)
class Material:
    """Isotropic-or-diagonal material (src/magnelio/materials/material.py:18-145)."""

    name: str = "air"
    epsilon: tuple[float, float, float] = (1.0, 1.0, 1.0)
    mu: tuple[float, float, float] = (1.0, 1.0, 1.0)
    sigma: tuple[float, float, float] = (0.0, 0.0, 0.0)
    sigma_m: tuple[float, float, float] = (0.0, 0.0, 0.0)
    is_pec: bool = False
    # pole-residue ADE model: list of (pole a_p, residue c_p) complex pairs in
    # rad/s (conjugate partners implied), eps(w) = eps_inf + sum c/(jw - a) + cc
    poles: tuple = ()
    # the same for the permeability, mu(w) = mu_inf + sum c/(jw - a) + cc, with
    # ``mu`` = mu_inf
    poles_mu: tuple = ()

    @classmethod
    def air(cls) -> "Material":
        return cls(name="air")

    @classmethod
    def pec(cls) -> "Material":
        return cls(name="PEC", is_pec=True)

    @classmethod
    def lossy_metal(cls, name: str, sigma: float, mu: float = 1.0) -> "Material":
        """A good conductor: PEC in the field solve, its finite sigma and mu
        used only by wall-loss models (surface impedance walls,
        src/magnelio/materials/material.py:148-185)."""
        if not 0.0 < sigma < float("inf"):
            raise ValueError(f"lossy_metal needs a finite sigma > 0, got {sigma!r}")
        return cls(name=name, mu=(mu,) * 3, sigma=(sigma,) * 3, is_pec=True)

    @classmethod
    def dielectric(cls, eps_r: float, sigma: float = 0.0, name: str = "dielectric") -> "Material":
        return cls(name=name, epsilon=(eps_r,) * 3, sigma=(sigma,) * 3)


@dataclass(frozen=True)
class GridLines:
    """Sorted node coordinates of a non-uniform hexahedral grid [m]."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    def __post_init__(self) -> None:
        for name in ("x", "y", "z"):
            arr = np.asarray(getattr(self, name), dtype=float)
            if arr.ndim != 1 or arr.size < 2:
                raise ValueError(f"GridLines.{name} must be a 1D array with at least 2 elements")
            if not np.all(np.diff(arr) > 0):
                raise ValueError(f"GridLines.{name} must be strictly increasing")
            object.__setattr__(self, name, arr)

    @property
    def Nx(self) -> int:
        return self.x.size - 1

    @property
    def Ny(self) -> int:
        return self.y.size - 1

    @property
    def Nz(self) -> int:
        return self.z.size - 1

    @property
    def dims(self) -> tuple[int, int, int]:
        return self.Nx, self.Ny, self.Nz

    @property
    def dx(self) -> np.ndarray:
        return np.diff(self.x)

    @property
    def dy(self) -> np.ndarray:
        return np.diff(self.y)

    @property
    def dz(self) -> np.ndarray:
        return np.diff(self.z)

    @property
    def n_cells(self) -> int:
        return self.Nx * self.Ny * self.Nz

    def centres(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        return (0.5 * (self.x[:-1] + self.x[1:]), 0.5 * (self.y[:-1] + self.y[1:]),
                0.5 * (self.z[:-1] + self.z[1:]))


def courant_dt(grid: GridLines, accuracy: str = "normal") -> float:
    """Geometric CFL step: safety / (c0 sqrt(1/dx_min^2 + 1/dy_min^2 + 1/dz_min^2))."""
    if accuracy not in SAFETY_FACTORS:
        raise ValueError(f"accuracy must be one of {list(SAFETY_FACTORS)!r}; got {accuracy!r}")
    dx, dy, dz = float(grid.dx.min()), float(grid.dy.min()), float(grid.dz.min())
    return SAFETY_FACTORS[accuracy] * (1.0 / (C0 * math.sqrt(1 / dx**2 + 1 / dy**2 + 1 / dz**2)))


@dataclass
class Mesh:
    """Grid + per-cell material ids + PEC edge mask (reference ``Mesh`` subset)."""

    grid: GridLines
    material_id: np.ndarray  # (Nx, Ny, Nz) int32
    material_library: dict[int, Material]
    pec_mask_edges: np.ndarray  # bool (3, n_max)
    pec_faces: tuple[str, ...] = field(default_factory=tuple)
    # pec_mask_edges is exactly the OR rule over PEC cells plus the pec_faces
    # walls (what the device coefficient builder recomputes); clear it when
    # the mask is edited by hand
    pec_rule: bool = True

    @property
    def dims(self) -> tuple[int, int, int]:
        return self.grid.dims

    @classmethod
    def from_grid(cls, grid: GridLines, regions=None, background: Material | None = None,
                  pec_faces=FACES, shapes=None) -> "Mesh":
        """AABB fill by cell centre (later regions win), then PEC mask + walls.

        ``regions``: [(Material, (xmin, ymin, zmin, xmax, ymax, zmax))], the
        reference's half-open ``lo <= c < hi`` cell-centre test.  ``shapes``:
        extra [(Material, predicate(xc, yc, zc) -> bool array)] applied after
        the boxes, in order (spheres, cylinders; this build's own fill).
        ``pec_faces``: bbox faces closed with PEC.
        """
        background = background or Material.air()
        nx, ny, nz = grid.dims
        lib: dict[int, Material] = {0: background}
        ids: dict[int, int] = {id(background): 0}
        mat_id = np.zeros((nx, ny, nz), dtype=np.int32)
        xc, yc, zc = grid.centres()

        def mid_of(mat: Material) -> int:
            key = id(mat)
            if key not in ids:
                ids[key] = len(lib)
                lib[ids[key]] = mat
            return ids[key]

        for mat, (x0, y0, z0, x1, y1, z1) in regions or ():
            m = mid_of(mat)
            ix = np.nonzero((xc >= x0) & (xc < x1))[0]
            iy = np.nonzero((yc >= y0) & (yc < y1))[0]
            iz = np.nonzero((zc >= z0) & (zc < z1))[0]
            if ix.size and iy.size and iz.size:
                mat_id[np.ix_(ix, iy, iz)] = m
        if shapes:
            X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij", sparse=True)
            for mat, pred in shapes:
                m = mid_of(mat)
                mat_id[np.broadcast_to(pred(X, Y, Z), mat_id.shape)] = m
        pec = build_pec_mask(grid, mat_id, lib)
        or_bbox_pec_walls(pec, nx, ny, nz, pec_faces)
        return cls(grid=grid, material_id=mat_id, material_library=lib, pec_mask_edges=pec,
                   pec_faces=tuple(pec_faces))


def _clamp_idx(n: int) -> np.ndarray:
    """Node index 0..n -> owning cell index, clamped to 0..n-1."""
    return np.minimum(np.arange(n + 1), max(n - 1, 0))


def build_pec_mask(grid: GridLines, mat_id: np.ndarray, lib: dict[int, Material]) -> np.ndarray:
    """(3, n_max) bool: an edge is PEC when any of its (clamped) cells is PEC."""
    nx, ny, nz = grid.dims
    pec_ids = [m for m, mat in lib.items() if mat.is_pec]
    cell = np.isin(mat_id, pec_ids) if pec_ids else np.zeros(mat_id.shape, dtype=bool)
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    # lower-side neighbour cells (index-1, clamped at 0)
    ilo = np.maximum(np.arange(nx + 1) - 1, 0)
    jlo = np.maximum(np.arange(ny + 1) - 1, 0)
    klo = np.maximum(np.arange(nz + 1) - 1, 0)

    def any4(a, i_sets, j_sets, k_sets):
        out = None
        for ii in i_sets:
            for jj in j_sets:
                for kk in k_sets:
                    v = a[np.ix_(ii, jj, kk)]
                    out = v if out is None else (out | v)
        return out

    ar_x, ar_y, ar_z = np.arange(nx), np.arange(ny), np.arange(nz)
    ex = any4(cell, [ar_x], [jlo, jl], [klo, kl])
    ey = any4(cell, [ilo, il], [ar_y], [klo, kl])
    ez = any4(cell, [ilo, il], [jlo, jl], [ar_z])
    n_max = max(ex.size, ey.size, ez.size)
    mask = np.zeros((3, n_max), dtype=bool)
    mask[0, : ex.size] = ex.ravel()
    mask[1, : ey.size] = ey.ravel()
    mask[2, : ez.size] = ez.ravel()
    return mask


def _edge_views(mask_row: np.ndarray, comp: int, nx: int, ny: int, nz: int) -> np.ndarray:
    shp = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)][comp]
    return mask_row[: int(np.prod(shp))].reshape(shp)


def or_bbox_pec_walls(pec: np.ndarray, nx: int, ny: int, nz: int, faces) -> None:
    """Mark the tangential E edges of each listed bbox face as PEC (in place)."""
    spec = {
        "xmin": ((1, (0, 0)), (2, (0, 0))), "xmax": ((1, (0, nx)), (2, (0, nx))),
        "ymin": ((0, (1, 0)), (2, (1, 0))), "ymax": ((0, (1, ny)), (2, (1, ny))),
        "zmin": ((0, (2, 0)), (1, (2, 0))), "zmax": ((0, (2, nz)), (1, (2, nz))),
    }
    for f in faces:
        if f not in spec:
            raise ValueError(f"unknown bbox face: {f!r}")
        for comp, (axis, idx) in spec[f]:
            v = _edge_views(pec[comp], comp, nx, ny, nz)
            sl = [slice(None)] * 3
            sl[axis] = idx
            v[tuple(sl)] = True


def _avg_d(d: np.ndarray) -> np.ndarray:
    """Dual widths at nodes 0..N: d[0], (d[i-1]+d[i])/2, d[N-1]."""
    n = d.size
    out = np.empty(n + 1)
    out[0] = d[0]
    out[n] = d[n - 1]
    if n > 1:
        out[1:n] = 0.5 * (d[:-1] + d[1:])
    return out


def _table(lib: dict[int, Material], prop: str, comp: int, pec_value=None) -> np.ndarray:
    t = np.zeros(max(lib) + 1)
    for m, mat in lib.items():
        t[m] = pec_value if (pec_value is not None and mat.is_pec) else getattr(mat, prop)[comp]
    return t


def _edge_mass(mesh: Mesh, prop: str, scale: float, pec_value=None) -> np.ndarray:
    g = mesh.grid
    nx, ny, nz = g.dims
    dx, dy, dz = g.dx, g.dy, g.dz
    ax, ay, az = _avg_d(dx), _avg_d(dy), _avg_d(dz)
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    mid = mesh.material_id
    tx, ty, tz = (_table(mesh.material_library, prop, c, pec_value) for c in range(3))
    # (scale * property) * geometry: the reference's evaluation order
    mx = (scale * tx[mid[:, jl][:, :, kl]]) * (ay[None, :, None] * az[None, None, :] / dx[:, None, None])
    my = (scale * ty[mid[il][:, :, kl]]) * (ax[:, None, None] * az[None, None, :] / dy[None, :, None])
    mz = (scale * tz[mid[il][:, jl]]) * (ax[:, None, None] * ay[None, :, None] / dz[None, None, :])
    return np.concatenate([mx.ravel(), my.ravel(), mz.ravel()])


def _face_mass(mesh: Mesh, prop: str, scale: float, pec_value=None) -> np.ndarray:
    g = mesh.grid
    nx, ny, nz = g.dims
    dx, dy, dz = g.dx, g.dy, g.dz
    ax, ay, az = _avg_d(dx), _avg_d(dy), _avg_d(dz)
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    mid = mesh.material_id
    tx, ty, tz = (_table(mesh.material_library, prop, c, pec_value) for c in range(3))
    hx = (scale * tx[mid[il]]) * (dy[None, :, None] * dz[None, None, :] / ax[:, None, None])
    hy = (scale * ty[mid[:, jl]]) * (dx[:, None, None] * dz[None, None, :] / ay[None, :, None])
    hz = (scale * tz[mid[:, :, kl]]) * (dx[:, None, None] * dy[None, :, None] / az[None, None, :])
    return np.concatenate([hx.ravel(), hy.ravel(), hz.ravel()])


def build_m_eps(mesh: Mesh) -> np.ndarray:
    """M_eps = EPS0 eps_r A_dual / L_primal (material_matrices.py:334-464, cat 0),
    with the thin-wire category-1 edges of the mesh (thinwire.py)."""
    m = _edge_mass(mesh, "epsilon", EPS0)
    wc = getattr(mesh, "wire_corrections", None)
    if wc is not None and wc.edge_idx.size:
        from .thinwire import correct_m_eps  # noqa: PLC0415

        geom = np.zeros(m.size)
        geom[wc.edge_idx] = _edge_geom_at(mesh.grid, wc.edge_idx)
        correct_m_eps(m, mesh, geom)
    return m


def build_m_mu(mesh: Mesh) -> np.ndarray:
    """M_mu = MU0 mu_r A_primal / L_dual (material_matrices.py:494-613, cat 0),
    with the thin-wire category-2 faces of the mesh (thinwire.py)."""
    m = _face_mass(mesh, "mu", MU0)
    wc = getattr(mesh, "wire_corrections", None)
    if wc is not None and wc.face_idx.size:
        from .thinwire import correct_m_mu  # noqa: PLC0415

        mu_face, l_dual = np.ones(m.size), np.ones(m.size)
        mu_face[wc.face_idx], l_dual[wc.face_idx] = _face_mu_ldual_at(mesh, wc.face_idx)
        correct_m_mu(m, mesh, mu_face, l_dual)
    return m


def _edge_geom_at(grid: GridLines, idx: np.ndarray) -> np.ndarray:
    """A_dual / L_primal of E edges (reference flat indices), as _edge_mass forms it."""
    nx, ny, nz = grid.dims
    ax, ay, az = _avg_d(grid.dx), _avg_d(grid.dy), _avg_d(grid.dz)
    shp = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    off = np.cumsum([0] + [int(np.prod(s)) for s in shp])
    out = np.empty(idx.size)
    for c in range(3):
        sel = (idx >= off[c]) & (idx < off[c + 1])
        i, j, k = np.unravel_index(idx[sel] - off[c], shp[c])
        if c == 0:
            out[sel] = ay[j] * az[k] / grid.dx[i]
        elif c == 1:
            out[sel] = ax[i] * az[k] / grid.dy[j]
        else:
            out[sel] = ax[i] * ay[j] / grid.dz[k]
    return out


def _face_mu_ldual_at(mesh: Mesh, idx: np.ndarray):
    """(staircase mu_r of the owning cell, dual length) of H faces."""
    g = mesh.grid
    nx, ny, nz = g.dims
    ax, ay, az = _avg_d(g.dx), _avg_d(g.dy), _avg_d(g.dz)
    il, jl, kl = _clamp_idx(nx), _clamp_idx(ny), _clamp_idx(nz)
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    off = np.cumsum([0] + [int(np.prod(s)) for s in shp])
    mu, ld = np.empty(idx.size), np.empty(idx.size)
    mid = mesh.material_id
    for c in range(3):
        sel = (idx >= off[c]) & (idx < off[c + 1])
        i, j, k = np.unravel_index(idx[sel] - off[c], shp[c])
        t = _table(mesh.material_library, "mu", c)
        if c == 0:
            mu[sel], ld[sel] = t[mid[il[i], j, k]], ax[i]
        elif c == 1:
            mu[sel], ld[sel] = t[mid[i, jl[j], k]], ay[j]
        else:
            mu[sel], ld[sel] = t[mid[i, j, kl[k]]], az[k]
    return mu, ld


def build_m_sigma(mesh: Mesh) -> np.ndarray:
    """M_sigma = sigma A_dual / L_primal, PEC -> 0 (material_matrices.py:616-686)."""
    return _edge_mass(mesh, "sigma", 1.0, pec_value=0.0)


def build_m_sigma_m(mesh: Mesh) -> np.ndarray:
    """M_sigma* = sigma_m A_primal / L_dual, PEC -> 0 (material_matrices.py:688-757)."""
    return _face_mass(mesh, "sigma_m", 1.0, pec_value=0.0)


_FACE_TANGENTIAL = {  # face -> (E comps, axis, boundary index fn, interior fn)
    "xmin": ((1, 2), 0), "xmax": ((1, 2), 0),
    "ymin": ((0, 2), 1), "ymax": ((0, 2), 1),
    "zmin": ((0, 1), 2), "zmax": ((0, 1), 2),
}
_NORMAL_H = {"xmin": 0, "xmax": 0, "ymin": 1, "ymax": 1, "zmin": 2, "zmax": 2}


def _e_component_views(vec: np.ndarray, nx: int, ny: int, nz: int) -> list[np.ndarray]:
    shp = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    out, off = [], 0
    for s in shp:
        n = int(np.prod(s))
        out.append(vec[off : off + n].reshape(s))
        off += n
    return out


def _h_component_views(vec: np.ndarray, nx: int, ny: int, nz: int) -> list[np.ndarray]:
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    out, off = [], 0
    for s in shp:
        n = int(np.prod(s))
        out.append(vec[off : off + n].reshape(s))
        off += n
    return out


def _copy_slab(view: np.ndarray, axis: int, dst: int, src: int) -> None:
    d = [slice(None)] * 3
    s = [slice(None)] * 3
    d[axis], s[axis] = dst, src
    view[tuple(d)] = view[tuple(s)]


def _boundary_and_interior(face: str, n_axis: int) -> tuple[int, int]:
    return (0, 1) if face.endswith("min") else (n_axis, n_axis - 1)


def flatten_port_mass(m_eps: np.ndarray, dims, face: str) -> np.ndarray:
    """Port-plane tangential M_eps <- first interior slab (material_matrices.py:257-331)."""
    nx, ny, nz = dims
    out = np.array(m_eps, copy=True)
    comps, axis = _FACE_TANGENTIAL[face]
    b, i = _boundary_and_interior(face, dims[axis])
    views = _e_component_views(out, nx, ny, nz)
    for c in comps:
        _copy_slab(views[c], axis, b, i)
    return out


def flatten_port_mu(m_mu: np.ndarray, dims, face: str) -> np.ndarray:
    """Port-plane normal M_mu <- first interior slab (material_matrices.py:181-255)."""
    nx, ny, nz = dims
    out = np.array(m_mu, copy=True)
    axis = _NORMAL_H[face]
    b, i = _boundary_and_interior(face, dims[axis])
    _copy_slab(_h_component_views(out, nx, ny, nz)[axis], axis, b, i)
    return out


def flatten_port_pec(pec: np.ndarray, dims, face: str) -> np.ndarray:
    """Port-plane tangential PEC mask <- first interior slab (material_matrices.py:106-179)."""
    nx, ny, nz = dims
    out = np.array(pec, copy=True)
    comps, axis = _FACE_TANGENTIAL[face]
    b, i = _boundary_and_interior(face, dims[axis])
    for c in comps:
        _copy_slab(_edge_views(out[c], c, nx, ny, nz), axis, b, i)
    return out


def pec_flat(pec: np.ndarray, dims) -> np.ndarray:
    """(3, n_max) mask -> flat length-n_E mask [Ex|Ey|Ez] (fit_td.py:481-490)."""
    nx, ny, nz = dims
    n = [nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz]
    return np.concatenate([pec[0, : n[0]], pec[1, : n[1]], pec[2, : n[2]]])


@dataclass
class Coefficients:
    """float64 update coefficients and the mass diagonals they came from."""

    alpha_e: np.ndarray
    beta_e: np.ndarray
    alpha_h: np.ndarray
    beta_h: np.ndarray
    m_eps: np.ndarray
    m_mu: np.ndarray
    pec_e: np.ndarray  # flat bool, length n_E


def update_coefficients(m_eps, m_mu, m_sigma, m_sigma_m, dt: float, pec_e: np.ndarray,
                        w_e=None, w_h=None) -> Coefficients:
    """Semi-implicit alpha/beta in float64 (fit_td.py:375-461), PEC freeze (:499-513).

    alpha = (M - dt/2 sigma + W)/(M + dt/2 sigma + W), beta = dt/(M + dt/2 sigma + W);
    M = 0 -> alpha = 1, beta = 0; PEC edges -> alpha = beta = 0.  ``w_e``/``w_h``
    are the already-halved pole terms 0.5*dt*W of dispersive materials (or None).
    """
    live_e = m_eps > 0
    if w_e is not None:
        denom = np.where(live_e, m_eps + 0.5 * dt * m_sigma + w_e, 1.0)
        alpha_e = np.where(live_e, (m_eps - 0.5 * dt * m_sigma + w_e) / denom, 1.0)
    else:
        denom = np.where(live_e, m_eps + 0.5 * dt * m_sigma, 1.0)
        alpha_e = np.where(live_e, (m_eps - 0.5 * dt * m_sigma) / denom, 1.0)
    beta_e = np.where(live_e, dt / denom, 0.0)

    live_h = m_mu > 0
    sm = np.where(live_h, m_sigma_m, 0.0)
    if w_h is not None:
        wh = np.where(live_h, w_h, 0.0)
        denom_h = np.where(live_h, m_mu + 0.5 * dt * sm + wh, 1.0)
        alpha_h = np.where(live_h, (m_mu - 0.5 * dt * sm + wh) / denom_h, 1.0)
    else:
        denom_h = np.where(live_h, m_mu + 0.5 * dt * sm, 1.0)
        alpha_h = np.where(live_h, (m_mu - 0.5 * dt * sm) / denom_h, 1.0)
    beta_h = np.where(live_h, dt / denom_h, 0.0)

    alpha_e[pec_e] = 0.0
    beta_e[pec_e] = 0.0
    return Coefficients(alpha_e, beta_e, alpha_h, beta_h, m_eps, m_mu, pec_e)
