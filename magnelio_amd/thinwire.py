"""Thin-wire sub-cell model: PEC edge chain plus the (m, 1/m) material
correction of the encircling cells.

Reference: src/magnelio/mesh/_thin_wire.py — ``_ring_stencil`` :76-218,
``_collect_requests`` :258-320, ``_write_corrections`` :358-430,
``apply_thin_wire_path`` :432-446 and ``correct_thin_wire_materials``
:532-580 — applied at mesh-build time (src/magnelio/mesh/mesher.py:
1067-1094).  The wire's E edges are PEC; per axis-aligned segment the
(up to four) encircling H faces get M_mu scaled by
    m = ln(delta / a) / ln(delta / r0),   r0 = KAPPA0 * geometric mean of
        the ring's transverse extents,    KAPPA0 = e^(-gamma) / 2^(3/2),
and the two radial E edges co-located with each face get M_eps / m, so the
pair product M_eps M_mu (the wave speed along the wire) is unchanged while
the line presents L' = (mu / 2 pi) ln(delta / a).  Requests onto one face
or edge (corners, neighbouring wires) keep the smallest m; faces whose own
extent is within 0.1 log units of r0 keep the bare value.

The reference takes its wire paths from CAD curves (OCC); here a wire is a
polyline of grid nodes (each leg along one axis), which is the
``EdgePath`` its own tests build by hand.  The corrected masses are
evaluated in the reference's float64 operation order: a corrected face is
the category-2 value MU0 mu A_face_free / L_dual with
A_face_free = m M_mu L_dual / (MU0 mu), a corrected edge the category-1
value EPS0 (eps_eff / m) A_dual / L with eps_eff = M_eps / (EPS0 A_dual / L).
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np

KAPPA0 = math.exp(-0.5772156649015329) / (2.0 * math.sqrt(2.0))
_RADIUS_MAX_FRAC = 0.30
_RADIUS_WARN_FRAC = 0.20
_MIN_LOG_DEN = 0.1
_ANISO_WARN_RATIO = 1.5


@dataclass(frozen=True)
class ThinWire:
    """A wire of radius ``radius`` [m] along the polyline ``points`` (grid
    nodes, consecutive points differing along one axis)."""

    name: str
    points: tuple
    radius: float


@dataclass
class WireCorrections:
    """Per-face / per-edge factors (reference flat layout) held by a mesh."""

    face_idx: np.ndarray
    face_m: np.ndarray
    edge_idx: np.ndarray
    edge_m: np.ndarray


def _sizes(nx, ny, nz):
    n_ex, n_ey = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    n_hx, n_hy = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    return n_ex, n_ey, n_hx, n_hy


def _e_flat(c, i, j, k, nx, ny, nz):
    n_ex, n_ey, _, _ = _sizes(nx, ny, nz)
    if c == 0:
        return (i * (ny + 1) + j) * (nz + 1) + k
    if c == 1:
        return n_ex + (i * ny + j) * (nz + 1) + k
    return n_ex + n_ey + (i * (ny + 1) + j) * nz + k


def _h_flat(c, i, j, k, nx, ny, nz):
    _, _, n_hx, n_hy = _sizes(nx, ny, nz)
    if c == 0:
        return (i * ny + j) * nz + k
    if c == 1:
        return n_hx + (i * (ny + 1) + j) * nz + k
    return n_hx + n_hy + (i * ny + j) * (nz + 1) + k


def rasterize(wire: ThinWire, grid) -> list:
    """[(axis, (i, j, k)), ...]: the E edges of the polyline, in order."""
    lines = (grid.x, grid.y, grid.z)
    nodes = []
    for p in wire.points:
        idx = []
        for a in range(3):
            q = int(np.argmin(np.abs(lines[a] - p[a])))
            if abs(lines[a][q] - p[a]) > 1e-9 * max(1.0, abs(p[a])) + 1e-12:
                raise ValueError(f"ThinWire {wire.name!r}: point {p} is not on a grid node")
            idx.append(q)
        nodes.append(tuple(idx))
    path = []
    for a, b in zip(nodes, nodes[1:]):
        diff = [b[t] - a[t] for t in range(3)]
        axes = [t for t in range(3) if diff[t] != 0]
        if len(axes) != 1:
            raise ValueError(f"ThinWire {wire.name!r}: leg {a} -> {b} is not along one axis")
        ax = axes[0]
        step = 1 if diff[ax] > 0 else -1
        for s in range(abs(diff[ax])):
            ijk = list(a)
            ijk[ax] = a[ax] + s * step if step > 0 else a[ax] - s - 1
            path.append((ax, tuple(ijk)))
    return path


def ring_stencil(axis: int, ijk, grid) -> list:
    """[(face_flat, delta_t, (edge_flat, edge_flat)), ...] of one segment
    (_thin_wire.py:76-218): the encircling H faces, their transverse extent
    from the wire and the two co-located radial E edges."""
    i, j, k = ijk
    nx, ny, nz = grid.dims
    dx, dy, dz = grid.dx, grid.dy, grid.dz
    out = []
    if axis == 2:
        for jf in (j, j - 1):
            if 0 <= jf <= ny - 1:
                out.append((_h_flat(0, i, jf, k, nx, ny, nz), float(dy[jf]),
                            (_e_flat(1, i, jf, k, nx, ny, nz), _e_flat(1, i, jf, k + 1, nx, ny, nz))))
        for i_f in (i, i - 1):
            if 0 <= i_f <= nx - 1:
                out.append((_h_flat(1, i_f, j, k, nx, ny, nz), float(dx[i_f]),
                            (_e_flat(0, i_f, j, k, nx, ny, nz), _e_flat(0, i_f, j, k + 1, nx, ny, nz))))
    elif axis == 0:
        for jf in (j, j - 1):
            if 0 <= jf <= ny - 1:
                out.append((_h_flat(2, i, jf, k, nx, ny, nz), float(dy[jf]),
                            (_e_flat(1, i, jf, k, nx, ny, nz), _e_flat(1, i + 1, jf, k, nx, ny, nz))))
        for kf in (k, k - 1):
            if 0 <= kf <= nz - 1:
                out.append((_h_flat(1, i, j, kf, nx, ny, nz), float(dz[kf]),
                            (_e_flat(2, i, j, kf, nx, ny, nz), _e_flat(2, i + 1, j, kf, nx, ny, nz))))
    elif axis == 1:
        for i_f in (i, i - 1):
            if 0 <= i_f <= nx - 1:
                out.append((_h_flat(2, i_f, j, k, nx, ny, nz), float(dx[i_f]),
                            (_e_flat(0, i_f, j, k, nx, ny, nz), _e_flat(0, i_f, j + 1, k, nx, ny, nz))))
        for kf in (k, k - 1):
            if 0 <= kf <= nz - 1:
                out.append((_h_flat(0, i, j, kf, nx, ny, nz), float(dz[kf]),
                            (_e_flat(2, i, j, kf, nx, ny, nz), _e_flat(2, i, j + 1, kf, nx, ny, nz))))
    else:
        raise ValueError(f"axis must be 0, 1 or 2, got {axis!r}")
    return out


def _validate_radius(radius, d_min, name):
    label = f"ThinWire {name!r}"
    if radius >= _RADIUS_MAX_FRAC * d_min:
        raise ValueError(f"{label}: radius {radius:.3e} m >= {_RADIUS_MAX_FRAC}x the smallest "
                         f"transverse cell along the path ({d_min:.3e} m)")
    if radius > _RADIUS_WARN_FRAC * d_min:
        warnings.warn(f"{label}: radius {radius:.3e} m is fatter than the bare-grid equivalent "
                      "radius; the correction factor drops below 1 and the stable step shrinks",
                      UserWarning, stacklevel=3)


def collect_requests(path, radius: float, grid, name: str = "") -> tuple[dict, dict]:
    """(face -> m, edge -> m) of one wire (_thin_wire.py:258-320)."""
    face_m: dict[int, float] = {}
    edge_m: dict[int, float] = {}
    n_weak = 0
    max_aniso = 1.0
    for axis, ijk in path:
        ring = ring_stencil(axis, ijk, grid)
        if not ring:
            continue
        ds = [d for _, d, _ in ring]
        max_aniso = max(max_aniso, max(ds) / min(ds))
        d_bar = math.exp(sum(math.log(d) for _, d, _ in ring) / len(ring))
        r0 = KAPPA0 * d_bar
        for face, delta_t, edges in ring:
            den = math.log(delta_t / r0)
            if den <= _MIN_LOG_DEN:
                n_weak += 1
                continue
            m = math.log(delta_t / radius) / den
            if face not in face_m or m < face_m[face]:
                face_m[face] = m
            for e in edges:
                if e not in edge_m or m < edge_m[e]:
                    edge_m[e] = m
    if n_weak:
        warnings.warn(f"ThinWire {name!r}: {n_weak} encircling face(s) keep the uncorrected value "
                      "(extreme grading)", UserWarning, stacklevel=3)
    if max_aniso > _ANISO_WARN_RATIO:
        warnings.warn(f"ThinWire {name!r}: anisotropic transverse cells at the wire (aspect up to "
                      f"{max_aniso:.2f}); the correction assumes locally isotropic cells",
                      UserWarning, stacklevel=3)
    return face_m, edge_m


def apply_thin_wires(mesh, wires, *, correction: bool = True) -> None:
    """Mask every wire's edges PEC and record the collapsed corrections on
    the mesh (mask_thin_wires + correct_thin_wire_materials).  The mesh's
    PEC mask is then no longer the plain OR rule (``pec_rule`` = False)."""
    from . import mesh as _mesh  # noqa: PLC0415

    grid = mesh.grid
    nx, ny, nz = grid.dims
    n_ex, n_ey, _, _ = _sizes(nx, ny, nz)
    face_m: dict[int, float] = {}
    edge_m: dict[int, float] = {}
    owner: dict[int, int] = {}
    n_shared = 0
    for w_i, wire in enumerate(wires):
        path = rasterize(wire, grid)
        d_min = min((d for ax, ijk in path for _, d, _ in ring_stencil(ax, ijk, grid)), default=math.inf)
        _validate_radius(wire.radius, d_min, wire.name)
        for ax, (i, j, k) in path:
            flat = _e_flat(ax, i, j, k, nx, ny, nz)
            local = flat - (0, n_ex, n_ex + n_ey)[ax]
            mesh.pec_mask_edges[ax, local] = True
        if not correction:
            continue
        f_m, e_m = collect_requests(path, wire.radius, grid, wire.name)
        for f, m in f_m.items():
            if f in face_m:
                n_shared += owner[f] != w_i
                face_m[f] = min(face_m[f], m)
            else:
                face_m[f], owner[f] = m, w_i
        for e, m in e_m.items():
            edge_m[e] = min(m, edge_m.get(e, math.inf))
    if n_shared:
        warnings.warn(f"{n_shared} encircling face(s) are shared between thin wires; the correction "
                      "keeps the smaller factor there", UserWarning, stacklevel=2)
    mesh.pec_rule = False
    pec_flat = _mesh.pec_flat(mesh.pec_mask_edges, grid.dims)
    edges = [(e, m) for e, m in sorted(edge_m.items()) if not pec_flat[e]]
    faces = sorted(face_m.items())
    prev = getattr(mesh, "wire_corrections", None)
    if prev is not None and (prev.face_idx.size or prev.edge_idx.size):
        raise ValueError("thin wires are applied to a mesh once (all wires in one call)")
    mesh.wire_corrections = WireCorrections(
        np.array([f for f, _ in faces], np.int64), np.array([m for _, m in faces], float),
        np.array([e for e, _ in edges], np.int64), np.array([m for _, m in edges], float))


def correct_m_mu(m_mu: np.ndarray, mesh, mu_face: np.ndarray, l_dual: np.ndarray) -> None:
    """Category-2 value of the corrected faces (material_matrices.py:591), in place."""
    from .mesh import MU0  # noqa: PLC0415

    wc = getattr(mesh, "wire_corrections", None)
    if wc is None or not wc.face_idx.size:
        return
    f = wc.face_idx
    a_free = ((wc.face_m * m_mu[f]) * l_dual[f]) / (MU0 * mu_face[f])
    m_mu[f] = ((MU0 * mu_face[f]) * a_free) / l_dual[f]


def correct_m_eps(m_eps: np.ndarray, mesh, geom_e: np.ndarray) -> None:
    """Category-1 value of the corrected edges (material_matrices.py:406), in place."""
    from .mesh import EPS0  # noqa: PLC0415

    wc = getattr(mesh, "wire_corrections", None)
    if wc is None or not wc.edge_idx.size:
        return
    e = wc.edge_idx
    eps_eff = m_eps[e] / (EPS0 * geom_e[e])
    m_eps[e] = (EPS0 * (eps_eff / wc.edge_m)) * geom_e[e]
