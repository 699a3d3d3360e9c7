"""Oracle: grid, AABB material fill, PEC mask, mass matrices, alpha/beta.

Restates (paths relative to the reference repository root):

* AABB fill           src/magnelio/mesh/mesher.py:255-279 (cell centre in
                      [lo, hi), later regions overwrite earlier ones)
* PEC edge mask       src/magnelio/mesh/indexing.py:168-257 — an edge is PEC
                      when ANY of its (index-clamped) adjacent cells is PEC
* bbox PEC walls      src/magnelio/mesh/mesher.py:1370-1430
* dual widths         src/magnelio/_operators/material_matrices.py:55-73
* M_eps/M_mu/M_sigma  material_matrices.py:334-464, :494-613, :616-686,
                      :688-757, category 0 (one clamped owning cell per
                      edge / face, EPS0*eps_r*A_dual/L etc.)
* port flattening     material_matrices.py:106-179, :181-255, :257-331
* alpha/beta          src/magnelio/solver/fit_td.py:375-513
* courant_dt          src/magnelio/solver/stability.py:44-93

This restatement works per element from explicit (i, j, k) index grids
(np.indices) rather than the reference's broadcast slices; the arithmetic
order of every product is kept, so values agree bit for bit.

Parity pins (tests/test_oracle_mesh.py): brute-force per-edge loops for the
PEC rule on small random meshes, closed forms on uniform grids
(M_eps = EPS0 eps_r dy dz / dx, boundary duals = full cell width), the
lossless limit alpha = 1, beta = dt/M, and the courant_dt closed form.
"""

from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0 * C0)


def courant_dt(x, y, z, safety: float = 0.95) -> float:
    dx, dy, dz = (float(np.diff(a).min()) for a in (x, y, z))
    return safety * (1.0 / (C0 * math.sqrt(1 / dx**2 + 1 / dy**2 + 1 / dz**2)))


def fill_aabb(x, y, z, boxes) -> np.ndarray:
    """material id per cell; boxes = [(mat_id, (x0, y0, z0, x1, y1, z1))]."""
    xc, yc, zc = (0.5 * (a[:-1] + a[1:]) for a in (x, y, z))
    mid = np.zeros((xc.size, yc.size, zc.size), dtype=np.int32)
    for m, (x0, y0, z0, x1, y1, z1) in boxes:
        inside = (
            ((xc >= x0) & (xc < x1))[:, None, None]
            & ((yc >= y0) & (yc < y1))[None, :, None]
            & ((zc >= z0) & (zc < z1))[None, None, :]
        )
        mid[inside] = m
    return mid


def dual_widths(d: np.ndarray) -> np.ndarray:
    n = d.size
    return np.array([d[0]] + [0.5 * (d[i - 1] + d[i]) for i in range(1, n)] + [d[n - 1]])


def _cl(a, n):
    return np.clip(a, 0, n - 1)


def pec_edge_mask(pec_cell: np.ndarray) -> np.ndarray:
    """Flat length-n_E bool mask [Ex|Ey|Ez] from a per-cell PEC flag."""
    nx, ny, nz = pec_cell.shape
    out = []
    # Ex (i, j, k): cells (i, j-1|j, k-1|k)
    i, j, k = np.indices((nx, ny + 1, nz + 1))
    m = np.zeros(i.shape, dtype=bool)
    for dj in (-1, 0):
        for dk in (-1, 0):
            m |= pec_cell[i, _cl(j + dj, ny), _cl(k + dk, nz)]
    out.append(m.ravel())
    i, j, k = np.indices((nx + 1, ny, nz + 1))
    m = np.zeros(i.shape, dtype=bool)
    for di in (-1, 0):
        for dk in (-1, 0):
            m |= pec_cell[_cl(i + di, nx), j, _cl(k + dk, nz)]
    out.append(m.ravel())
    i, j, k = np.indices((nx + 1, ny + 1, nz))
    m = np.zeros(i.shape, dtype=bool)
    for di in (-1, 0):
        for dj in (-1, 0):
            m |= pec_cell[_cl(i + di, nx), _cl(j + dj, ny), k]
    out.append(m.ravel())
    return np.concatenate(out)


def bbox_walls(pec_flat: np.ndarray, nx, ny, nz, faces) -> np.ndarray:
    """OR the tangential E edges of the listed faces into a flat mask."""
    out = pec_flat.copy()
    sh = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    offs = np.cumsum([0] + [int(np.prod(s)) for s in sh])
    comps = [out[offs[c] : offs[c + 1]].reshape(sh[c]) for c in range(3)]
    for f in faces:
        axis = "xyz".index(f[0])
        idx = 0 if f.endswith("min") else (nx, ny, nz)[axis]
        for c in range(3):
            if c == axis:
                continue
            sl = [slice(None)] * 3
            sl[axis] = idx
            comps[c][tuple(sl)] = True
    return out


def mass_e(x, y, z, mat_id, table, scale):
    """Per-edge scale*prop(owning cell)*A_dual/L_primal, flat [Ex|Ey|Ez].

    ``table[c][m]`` is component c of the material property of id m.
    """
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    nx, ny, nz = dx.size, dy.size, dz.size
    ax, ay, az = dual_widths(dx), dual_widths(dy), dual_widths(dz)
    out = []
    i, j, k = np.indices((nx, ny + 1, nz + 1))
    p = np.asarray(table[0])[mat_id[i, _cl(j, ny), _cl(k, nz)]]
    out.append(((scale * p) * (ay[j] * az[k] / dx[i])).ravel())
    i, j, k = np.indices((nx + 1, ny, nz + 1))
    p = np.asarray(table[1])[mat_id[_cl(i, nx), j, _cl(k, nz)]]
    out.append(((scale * p) * (ax[i] * az[k] / dy[j])).ravel())
    i, j, k = np.indices((nx + 1, ny + 1, nz))
    p = np.asarray(table[2])[mat_id[_cl(i, nx), _cl(j, ny), k]]
    out.append(((scale * p) * (ax[i] * ay[j] / dz[k])).ravel())
    return np.concatenate(out)


def mass_h(x, y, z, mat_id, table, scale):
    """Per-face scale*prop(owning cell)*A_primal/L_dual, flat [Hx|Hy|Hz]."""
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    nx, ny, nz = dx.size, dy.size, dz.size
    ax, ay, az = dual_widths(dx), dual_widths(dy), dual_widths(dz)
    out = []
    i, j, k = np.indices((nx + 1, ny, nz))
    p = np.asarray(table[0])[mat_id[_cl(i, nx), j, k]]
    out.append(((scale * p) * (dy[j] * dz[k] / ax[i])).ravel())
    i, j, k = np.indices((nx, ny + 1, nz))
    p = np.asarray(table[1])[mat_id[i, _cl(j, ny), k]]
    out.append(((scale * p) * (dx[i] * dz[k] / ay[j])).ravel())
    i, j, k = np.indices((nx, ny, nz + 1))
    p = np.asarray(table[2])[mat_id[i, j, _cl(k, nz)]]
    out.append(((scale * p) * (dx[i] * dy[j] / az[k])).ravel())
    return np.concatenate(out)


def flatten_port(vec: np.ndarray, nx, ny, nz, face: str, side: str) -> np.ndarray:
    """Copy the first interior slab onto the port-plane slab.

    side='E': the two tangential E components (M_eps, PEC mask);
    side='H': the normal H component (M_mu).
    """
    out = vec.copy()
    axis = "xyz".index(face[0])
    if side == "E":
        sh = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
        comps = [c for c in range(3) if c != axis]
    else:
        sh = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
        comps = [axis]
    offs = np.cumsum([0] + [int(np.prod(s)) for s in sh])
    n_ax = (nx, ny, nz)[axis]
    b, src = (0, 1) if face.endswith("min") else (n_ax, n_ax - 1)
    for c in comps:
        v = out[offs[c] : offs[c + 1]].reshape(sh[c])
        dst_sl = [slice(None)] * 3
        src_sl = [slice(None)] * 3
        dst_sl[axis], src_sl[axis] = b, src
        v[tuple(dst_sl)] = v[tuple(src_sl)]
    return out


def alpha_beta(m, sigma, dt, frozen_mask=None, w=None):
    """alpha = (M - dt/2 s + W)/(M + dt/2 s + W), beta = dt/(...); M <= 0 frozen
    (alpha 1, beta 0); ``frozen_mask`` (PEC) -> alpha = beta = 0.
    ``w`` is the already halved dispersive term 0.5*dt*W (or None)."""
    live = m > 0
    num = m - 0.5 * dt * sigma
    den = m + 0.5 * dt * sigma
    if w is not None:
        num = num + w
        den = den + w
    den = np.where(live, den, 1.0)
    alpha = np.where(live, num / den, 1.0)
    beta = np.where(live, dt / den, 0.0)
    if frozen_mask is not None:
        alpha = np.where(frozen_mask, 0.0, alpha)
        beta = np.where(frozen_mask, 0.0, beta)
    return alpha, beta
