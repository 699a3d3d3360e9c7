"""Modal waveguide ports: plane maps, discrete TE modes, calibration, DTBC setup.

Host-side float64 setup of the modal port operator whose per-step work
(interior projection, DTBC recursion, modal write-back, V/I recording) runs
on the device in ``csrc/mgh_ports.hip``.

Reference counterparts (relative to the reference repository root):

* ``PortPlane.from_mesh``          src/magnelio/ports/_modal/port_plane.py:32-334
                                   (whole-face planes; (u, v) per face from
                                   src/magnelio/mesh/faces.py:63-71)
* 2-D curl-curl restriction        ports/_modal/curl_curl_2d.py:52-124, :416-471
* numerical TE modes (eigsh)       ports/_modal/numerical_2d.py:454-583, sign
                                   convention :758-802, Z at f_calc :828-853
* numerical TM modes               normal-E node problem ports/_modal/
                                   curl_curl_2d.py:127-218 (normal edges
                                   :236-292, plane gradient :295-347), eigsh
                                   numerical_2d.py:585-753; TE + TM merge by
                                   cut-off ports/_modal/factory.py:1627-1660
* travelling-wave H profiles       ports/_modal/tem_laplace.py:102-167
* V/I calibration, physical scales ports/_modal/operator.py:778-941 (numerical
                                   path; staircase planes: flux factor 1)
* DTBC chain certificate, z0       ports/_modal/operator.py:671-776
* slab-defect certificate          ports/_modal/factory.py:1189-1244
* Mur coefficient / delay          ports/_modal/operator.py:652-661
* complement absorber arrays       ports/_modal/factory.py:687-735
* feed validation                  ports/_modal/factory.py:875-950
* build order                      ports/_modal/factory.py:1367-1877

Scope: whole-face ports on hollow (single-dielectric) feeds; the channels
are the lowest ``n_modes`` discrete cut-offs over the TE and TM families
together (TEM / quasi-TEM multi-conductor feeds are not built).
Degenerate-mode gauge fixing is not built: a degenerate pair keeps the
eigen-solver's basis.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh

from . import mesh as _mesh
from .dtbc import N_KERNEL_INIT, packed_kernels

C0, MU0, EPS0 = _mesh.C0, _mesh.MU0, _mesh.EPS0
ETA0 = math.sqrt(MU0 / EPS0)

# face -> (normal axis, is_max, u_axis, v_axis); u x v points into the domain
FACES = {
    "xmin": (0, False, 1, 2), "xmax": (0, True, 2, 1),
    "ymin": (1, False, 2, 0), "ymax": (1, True, 0, 2),
    "zmin": (2, False, 0, 1), "zmax": (2, True, 1, 0),
}

DTBC_PAIR_SPREAD_TOL = 1e-8
DTBC_WEIGHT_FLOOR = 1e-12
DTBC_SLAB_DEFECT_TOL = 1e-8
DEGENERATE_CUTOFF_RTOL = 1e-3


def _e_shape(c, nx, ny, nz):
    return [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)][c]


def _h_shape(c, nx, ny, nz):
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)][c]


def _offsets(shapes):
    out, o = [], 0
    for s in shapes:
        out.append(o)
        o += int(np.prod(s))
    return out


@dataclass(frozen=True)
class PortPlane:
    """Tangential edge maps of one whole bbox face (reference flat layout)."""

    face: str
    coordinate: float
    e_u_indices: np.ndarray
    h_v_indices: np.ndarray
    u_edge_uv: np.ndarray
    u_edge_lengths: np.ndarray
    e_v_indices: np.ndarray
    h_u_indices: np.ndarray
    v_edge_uv: np.ndarray
    v_edge_lengths: np.ndarray
    e_u_indices_interior: np.ndarray
    e_v_indices_interior: np.ndarray
    normal_dx: float
    u_node_window: tuple[int, int]
    v_node_window: tuple[int, int]

    @property
    def normal_axis(self) -> int:
        return FACES[self.face][0]

    @property
    def is_max(self) -> bool:
        return FACES[self.face][1]

    @property
    def u_axis(self) -> int:
        return FACES[self.face][2]

    @property
    def v_axis(self) -> int:
        return FACES[self.face][3]

    @classmethod
    def from_mesh(cls, face: str, grid) -> "PortPlane":
        if face not in FACES:
            raise ValueError(f"unknown face {face!r}")
        n_ax, is_max, u_ax, v_ax = FACES[face]
        nx, ny, nz = grid.dims
        N = (nx, ny, nz)
        nodes = (grid.x, grid.y, grid.z)
        cent = grid.centres()
        delt = (grid.dx, grid.dy, grid.dz)
        e_off = _offsets([_e_shape(c, nx, ny, nz) for c in range(3)])
        h_off = _offsets([_h_shape(c, nx, ny, nz) for c in range(3)])
        k_port = N[n_ax] if is_max else 0
        k_int = N[n_ax] - 1 if is_max else 1
        k_h = N[n_ax] - 1 if is_max else 0
        normal_dx = float(delt[n_ax][k_h])

        def edges(p_ax, s_ax, kn_e):
            # edges along p_ax on cells [0, N_p), at secondary nodes [0, N_s]
            pi, si = np.meshgrid(np.arange(N[p_ax]), np.arange(N[s_ax] + 1), indexing="ij")
            ijk_e = [None, None, None]
            ijk_e[n_ax], ijk_e[p_ax], ijk_e[s_ax] = np.full_like(pi, kn_e), pi, si
            es = _e_shape(p_ax, nx, ny, nz)
            e_idx = (e_off[p_ax] + (ijk_e[0] * es[1] + ijk_e[1]) * es[2] + ijk_e[2]).ravel()
            ijk_h = [None, None, None]
            ijk_h[n_ax], ijk_h[p_ax], ijk_h[s_ax] = np.full_like(pi, k_h), pi, si
            hs = _h_shape(s_ax, nx, ny, nz)
            h_idx = (h_off[s_ax] + (ijk_h[0] * hs[1] + ijk_h[1]) * hs[2] + ijk_h[2]).ravel()
            uv = np.stack([cent[p_ax][pi].ravel(), nodes[s_ax][si].ravel()], axis=1)
            return (e_idx.astype(np.int64), h_idx.astype(np.int64), uv, delt[p_ax][pi].ravel())

        eu, hv, uuv, ulen = edges(u_ax, v_ax, k_port)
        eu_i, _, _, _ = edges(u_ax, v_ax, k_int)
        ev, hu, vuv_raw, vlen = edges(v_ax, u_ax, k_port)
        ev_i, _, _, _ = edges(v_ax, u_ax, k_int)
        return cls(face=face, coordinate=float(nodes[n_ax][k_port]), e_u_indices=eu,
                   h_v_indices=hv, u_edge_uv=uuv, u_edge_lengths=ulen, e_v_indices=ev,
                   h_u_indices=hu, v_edge_uv=vuv_raw[:, [1, 0]], v_edge_lengths=vlen,
                   e_u_indices_interior=eu_i, e_v_indices_interior=ev_i, normal_dx=normal_dx,
                   u_node_window=(0, N[u_ax]), v_node_window=(0, N[v_ax]))


def validate_feed(face: str, grid, rtol: float = 1e-6) -> None:
    """The three cells next to the port plane must be equidistant."""
    n_ax, is_max, _, _ = FACES[face]
    d = (grid.dx, grid.dy, grid.dz)[n_ax]
    if d.size < 3:
        raise ValueError(f"port {face!r}: at least 3 cells required along the port-normal axis")
    cells = d[-3:] if is_max else d[:3]
    cmin, cmax = float(cells.min()), float(cells.max())
    if (cmax - cmin) / cmin > rtol:
        raise ValueError(f"port {face!r}: the three port-adjacent cells must be equidistant "
                         f"(got {cells.tolist()})")


def normal_h_indices(face: str, grid) -> np.ndarray:
    """Flat H indices of the normal H faces on the port plane."""
    n_ax, is_max, _, _ = FACES[face]
    nx, ny, nz = grid.dims
    h_off = _offsets([_h_shape(c, nx, ny, nz) for c in range(3)])
    hs = _h_shape(n_ax, nx, ny, nz)
    kn = (nx, ny, nz)[n_ax] if is_max else 0
    t = [a for a in range(3) if a != n_ax]
    a_i, b_i = np.meshgrid(np.arange((nx, ny, nz)[t[0]]), np.arange((nx, ny, nz)[t[1]]), indexing="ij")
    ijk = [None, None, None]
    ijk[n_ax], ijk[t[0]], ijk[t[1]] = np.full_like(a_i, kn), a_i, b_i
    return (h_off[n_ax] + (ijk[0] * hs[1] + ijk[1]) * hs[2] + ijk[2]).ravel().astype(np.int64)


def curl_2d(plane: PortPlane, grid) -> sp.csr_matrix:
    """Rows: normal H faces of the plane; columns: [e_u | e_v] edges.

    Circulation of the four tangential edges around each normal face,
    curl_n = dE_q/dp - dE_p/dq with (n, p, q) cyclic.
    """
    nx, ny, nz = grid.dims
    n_ax = plane.normal_axis
    p_ax, q_ax = (n_ax + 1) % 3, (n_ax + 2) % 3
    e_off = _offsets([_e_shape(c, nx, ny, nz) for c in range(3)])
    kn = (nx, ny, nz)[n_ax] if plane.is_max else 0
    prim = np.concatenate([plane.e_u_indices, plane.e_v_indices])
    order = np.argsort(prim)
    sprim = prim[order]
    N = (nx, ny, nz)
    ap, aq = np.meshgrid(np.arange(N[p_ax]), np.arange(N[q_ax]), indexing="ij")
    ap, aq = ap.ravel(), aq.ravel()

    def flat(comp, ip, iq):
        ijk = [None, None, None]
        ijk[n_ax], ijk[p_ax], ijk[q_ax] = np.full_like(ip, kn), ip, iq
        es = _e_shape(comp, nx, ny, nz)
        return e_off[comp] + (ijk[0] * es[1] + ijk[1]) * es[2] + ijk[2]

    terms = [(flat(q_ax, ap + 1, aq), +1.0), (flat(q_ax, ap, aq), -1.0),
             (flat(p_ax, ap, aq + 1), -1.0), (flat(p_ax, ap, aq), +1.0)]
    rows, cols, vals = [], [], []
    r = np.arange(ap.size)
    for idx, sgn in terms:
        pos = order[np.searchsorted(sprim, idx)]
        rows.append(r)
        cols.append(pos)
        vals.append(np.full(r.size, sgn))
    # row order: the same raster as normal_h_indices (global tangential axes)
    c = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ap.size, prim.size))
    return c


def _arpack_v0(n: int) -> np.ndarray:
    return np.random.default_rng(0).standard_normal(n)


def z_te(omega: float, omega_c: float, eps_r: float) -> complex:
    """TE wave impedance j w mu0 / gamma (src/magnelio/ports/_modal/mode.py:168-213)."""
    return 1j * omega * MU0 / gamma_mode(omega, omega_c, eps_r)


def z_tm(omega: float, omega_c: float, eps_r: float) -> complex:
    """TM wave impedance gamma / (j w eps0 eps_r) (mode.py:206-211)."""
    return gamma_mode(omega, omega_c, eps_r) / (1j * omega * EPS0 * eps_r)


def gamma_mode(omega: float, omega_c: float, eps_r: float) -> complex:
    k2 = omega**2 * MU0 * EPS0 * eps_r
    kc2 = omega_c**2 * MU0 * EPS0 * eps_r
    diff = kc2 - k2
    return 1j * math.sqrt(-diff) if diff <= 0.0 else complex(math.sqrt(diff))


@dataclass
class DiscreteMode:
    """One discrete hollow-guide channel (TE or TM) on a port plane."""

    name: str
    omega_c: float
    epsilon_r: float
    e_u: np.ndarray
    e_v: np.ndarray
    h_u: np.ndarray
    h_v: np.ndarray
    mode_type: str = "TE"
    z_line: float | None = None  # TEM: the line impedance

    def z_modal(self, omega: float) -> complex:
        """Z of the power-wave split: the line impedance of a TEM channel,
        else the wave impedance (mode.py:215-231)."""
        if self.z_line is not None:
            return complex(self.z_line)
        if self.mode_type == "TEM":
            return complex(ETA0 / math.sqrt(self.epsilon_r))
        if self.mode_type == "TM":
            return z_tm(omega, self.omega_c, self.epsilon_r)
        return z_te(omega, self.omega_c, self.epsilon_r)

    def gamma(self, omega: float) -> complex:
        return gamma_mode(omega, self.omega_c, self.epsilon_r)


def _resolve_sign(plane: PortPlane, v: np.ndarray) -> np.ndarray:
    """Port-symmetric sign from global tangential coordinates (numerical_2d.py:758-802)."""
    if plane.u_axis < plane.v_axis:
        u_lo, u_hi = plane.u_edge_uv[:, 0], plane.u_edge_uv[:, 1]
        v_lo, v_hi = plane.v_edge_uv[:, 0], plane.v_edge_uv[:, 1]
    else:
        u_lo, u_hi = plane.u_edge_uv[:, 1], plane.u_edge_uv[:, 0]
        v_lo, v_hi = plane.v_edge_uv[:, 1], plane.v_edge_uv[:, 0]
    lo = np.concatenate([u_lo, v_lo])
    hi = np.concatenate([u_hi, v_hi])
    dlo = lo - 0.5 * (lo.max() + lo.min())
    dhi = hi - 0.5 * (hi.max() + hi.min())
    vmax = float(np.max(np.abs(v))) if v.size else 1.0
    for f in (dlo, dhi, dlo * dhi, dlo * dlo + dhi * dhi):
        fmax = float(np.max(np.abs(f))) if f.size else 1.0
        ov = float(np.sum(v * f))
        if abs(ov) > 1e-9 * vmax * fmax * max(v.size, 1):
            return -v if ov < 0.0 else v
    j = int(np.argmax(np.abs(v)))
    return -v if v[j] < 0.0 else v


def solve_te_modes(plane: PortPlane, grid, m_eps, m_mu, pec_e_flat, n_modes: int,
                   f_calc: float, eps_r: float = 1.0) -> list[DiscreteMode]:
    """Lowest TE modes of K e = w_c^2 M e on the plane (free, non-PEC edges)."""
    prim = np.concatenate([plane.e_u_indices, plane.e_v_indices])
    hn = normal_h_indices(plane.face, grid)
    C = curl_2d(plane, grid)
    mmu = m_mu[hn]
    mu_inv = np.where(mmu > 0, 1.0 / np.where(mmu > 0, mmu, 1.0), 0.0)
    K = (C.T @ sp.diags(mu_inv) @ C).tocsr()
    Md = m_eps[prim]
    pec = pec_e_flat[prim]
    if not pec.any():
        raise ValueError(f"port {plane.face}: the plane has no PEC wall edges (hollow feed required)")
    free = np.nonzero(~pec)[0]
    Kf = K[free][:, free]
    Mf = sp.diags(Md[free])
    L_u = float(plane.u_edge_uv[:, 0].max() - plane.u_edge_uv[:, 0].min())
    L_v = float(plane.v_edge_uv[:, 1].max() - plane.v_edge_uv[:, 1].min())
    c_eff = C0 / math.sqrt(eps_r)
    sigma = (0.95 * math.pi * c_eff / max(L_u, L_v)) ** 2
    nf = free.size
    if nf <= 600:
        from scipy.linalg import eigh  # noqa: PLC0415

        vals, vecs = eigh(Kf.toarray(), np.diag(Md[free]))
    else:
        n_total = min(max(n_modes + 8, 2 * n_modes), nf - 2)
        while True:
            vals, vecs = eigsh(Kf, k=n_total, M=Mf, sigma=sigma, v0=_arpack_v0(nf))
            if np.count_nonzero(vals > 1e-3 * sigma) >= n_modes or n_total >= nf - 2:
                break
            n_total = min(2 * n_total, nf - 2)
    phys = vals > 1e-3 * sigma
    vals, vecs = vals[phys], vecs[:, phys]
    order = np.argsort(vals)[:n_modes]
    if order.size < n_modes:
        raise RuntimeError(f"only {order.size} TE modes found on port {plane.face}")
    n_u = plane.e_u_indices.size
    omega_calc = 2 * math.pi * f_calc
    modes = []
    for rank, i in enumerate(order):
        full = np.zeros(prim.size)
        full[free] = vecs[:, i]
        full /= math.sqrt(float(np.sum(Md * full * full)))  # M_eps-orthonormal
        full = _resolve_sign(plane, full)
        e_u, e_v = full[:n_u].copy(), full[n_u:].copy()
        omega_c = math.sqrt(float(vals[i]))
        z_real = _z_real("TE", omega_calc, omega_c, eps_r)
        h_u, h_v = travelling_wave_h(e_u, e_v, plane, m_mu, z_real)
        modes.append(DiscreteMode(f"TE_num{rank:02d}", omega_c, eps_r, e_u, e_v, h_u, h_v))
    return modes


def normal_e_indices(plane: PortPlane, grid) -> np.ndarray:
    """Flat E indices of the normal edges leaving the plane's nodes into the
    domain (layer 0 on a MIN face, N - 1 on a MAX face), one per plane node
    in the (u, v) node raster iu * (N_v + 1) + iv (curl_curl_2d.py:236-292)."""
    n_ax, is_max, u_ax, v_ax = FACES[plane.face]
    nx, ny, nz = grid.dims
    N = (nx, ny, nz)
    e_off = _offsets([_e_shape(c, nx, ny, nz) for c in range(3)])
    iu, iv = np.meshgrid(np.arange(N[u_ax] + 1), np.arange(N[v_ax] + 1), indexing="ij")
    ijk = [None, None, None]
    ijk[n_ax], ijk[u_ax], ijk[v_ax] = np.full_like(iu, N[n_ax] - 1 if is_max else 0), iu, iv
    es = _e_shape(n_ax, nx, ny, nz)
    return (e_off[n_ax] + (ijk[0] * es[1] + ijk[1]) * es[2] + ijk[2]).ravel().astype(np.int64)


def gradient_2d(plane: PortPlane) -> sp.csr_matrix:
    """Plane-node gradient: rows [e_u | e_v] in plane order, columns the
    node raster of :func:`normal_e_indices`; +1 at the edge's end node (the
    restriction of the 3-D node gradient, curl_curl_2d.py:295-347)."""
    n_u_c, n_v_c = plane.u_node_window[1], plane.v_node_window[1]  # cells along u, v
    nv1 = n_v_c + 1
    # e_u (iu cell, iv node), iv fastest: nodes (iu, iv) -> (iu + 1, iv)
    iu, iv = np.meshgrid(np.arange(n_u_c), np.arange(n_v_c + 1), indexing="ij")
    a_u, b_u = (iu * nv1 + iv).ravel(), ((iu + 1) * nv1 + iv).ravel()
    # e_v (iv cell, iu node), iu fastest: nodes (iu, iv) -> (iu, iv + 1)
    jv, ju = np.meshgrid(np.arange(n_v_c), np.arange(n_u_c + 1), indexing="ij")
    a_v, b_v = (ju * nv1 + jv).ravel(), (ju * nv1 + jv + 1).ravel()
    a, b = np.concatenate([a_u, a_v]), np.concatenate([b_u, b_v])
    r = np.arange(a.size)
    return sp.csr_matrix((np.concatenate([-np.ones(a.size), np.ones(a.size)]),
                          (np.concatenate([r, r]), np.concatenate([a, b]))),
                         shape=(a.size, (n_u_c + 1) * nv1))


def solve_tm_modes(plane: PortPlane, grid, m_eps, m_mu, pec_e_flat, n_modes: int,
                   f_calc: float, eps_r: float = 1.0) -> list[DiscreteMode]:
    """Lowest TM modes of the normal-E node problem K_z phi = w_c^2 M_z phi.

    The normal edges at the plane exchange energy with the transversal H
    faces of the first cell only: K_z = C_s^T diag(1/M_mu) C_s with C_s the
    3-D curl restricted to those faces and edges, which on the plane is the
    node gradient, so K_z = G^T diag(1/M_mu[h co-located with each edge]) G;
    M_z = M_eps of the normal edges; Dirichlet where the normal edge is PEC
    or massless (curl_curl_2d.py:127-218).  Shift-invert eigsh below the
    hollow-rectangle TM11 cut-off (numerical_2d.py:585-753); the transverse
    profile is e_t = G phi, M_eps-normalised on the plane edges and signed
    by the port-symmetric rule.
    """
    if f_calc <= 0.0:
        raise ValueError("f_calc must be positive")
    e_n = normal_e_indices(plane, grid)
    G = gradient_2d(plane)
    mh = np.concatenate([m_mu[plane.h_v_indices], m_mu[plane.h_u_indices]])  # [e_u | e_v] rows
    mu_inv = np.where(mh > 0, 1.0 / np.where(mh > 0, mh, 1.0), 0.0)
    K = (G.T @ sp.diags(mu_inv) @ G).tocsr()
    m_z = np.asarray(m_eps[e_n], float)
    dirichlet = np.asarray(pec_e_flat[e_n], bool) | (m_z <= 0.0)
    free = np.nonzero(~dirichlet)[0]
    if free.size == 0:
        raise RuntimeError(f"port {plane.face}: every plane node is PEC; no TM degrees of freedom")
    Kf = K[free][:, free]
    Mf = sp.diags(m_z[free]).tocsr()
    L_u = float(plane.u_edge_uv[:, 0].max() - plane.u_edge_uv[:, 0].min())
    L_v = float(plane.v_edge_uv[:, 1].max() - plane.v_edge_uv[:, 1].min())
    if L_u <= 0.0 or L_v <= 0.0:
        raise RuntimeError(f"port {plane.face}: zero u/v extent")
    c_eff = C0 / math.sqrt(eps_r)
    sigma = 0.95 * 0.95 * ((math.pi * c_eff / L_u) ** 2 + (math.pi * c_eff / L_v) ** 2)
    n_total = min(max(n_modes + 4, 2 * n_modes), free.size - 2)
    if n_total < n_modes:
        raise RuntimeError(f"port {plane.face}: {free.size} free TM nodes are too few for "
                           f"{n_modes} modes")
    vals, vecs = eigsh(Kf, k=n_total, M=Mf, sigma=sigma, v0=_arpack_v0(free.size))
    if np.any(vals <= 0.0):
        raise RuntimeError(f"port {plane.face}: non-positive TM eigenvalues {vals[vals <= 0.0]}")
    order = np.argsort(vals)[:n_modes]
    phi = np.zeros((e_n.size, order.size))
    phi[free] = vecs[:, order]
    e_t = G @ phi
    me2d = m_eps[np.concatenate([plane.e_u_indices, plane.e_v_indices])]
    norm_sq = np.einsum("ij,i,ij->j", e_t, me2d, e_t)
    if np.any(norm_sq <= 0.0):
        raise RuntimeError(f"port {plane.face}: TM profile with a non-positive M_eps norm")
    e_t = e_t / np.sqrt(norm_sq)[None, :]
    n_u = plane.e_u_indices.size
    omega_calc = 2 * math.pi * f_calc
    modes = []
    for rank in range(order.size):
        full = _resolve_sign(plane, e_t[:, rank])
        e_u, e_v = full[:n_u].copy(), full[n_u:].copy()
        omega_c = math.sqrt(float(vals[order[rank]]))
        h_u, h_v = travelling_wave_h(e_u, e_v, plane, m_mu, _z_real("TM", omega_calc, omega_c, eps_r))
        modes.append(DiscreteMode(f"TM_num{rank:02d}", omega_c, eps_r, e_u, e_v, h_u, h_v, "TM"))
    return modes


def solve_modes(plane: PortPlane, grid, m_eps, m_mu, pec_e_flat, n_modes: int, f_calc: float,
                eps_r: float = 1.0) -> list[DiscreteMode]:
    """The n_modes lowest cut-offs over both families: n_modes of each are
    solved, merged by a stable sort on w_c (TE first on a tie) and cut
    (factory.py:1627-1660)."""
    te = solve_te_modes(plane, grid, m_eps, m_mu, pec_e_flat, n_modes, f_calc, eps_r)
    tm = solve_tm_modes(plane, grid, m_eps, m_mu, pec_e_flat, n_modes, f_calc, eps_r)
    return sorted(te + tm, key=lambda m: m.omega_c)[:n_modes]


def _z_real(mode_type: str, omega: float, omega_c: float, eps_r: float) -> float:
    """|Z| at omega for the H-profile bake-in (numerical_2d.py:828-853)."""
    k2 = omega**2 * MU0 * EPS0 * eps_r
    kc2 = omega_c**2 * MU0 * EPS0 * eps_r
    diff = kc2 - k2
    if abs(diff) < 1e-30:
        raise ValueError("f_calc coincides with a mode cut-off")
    if mode_type == "TE":
        return omega * MU0 / math.sqrt(abs(diff))
    return math.sqrt(abs(diff)) / (omega * EPS0 * eps_r)


def travelling_wave_h(e_u, e_v, plane: PortPlane, m_mu, z_wave: float):
    """h_u = -e_v (MU0 ndx / Z)/M_mu, h_v = +e_u (MU0 ndx / Z)/M_mu (tem_laplace.py:102-167)."""
    scale = MU0 * float(plane.normal_dx) / float(z_wave)
    mu_u = m_mu[plane.h_u_indices]
    mu_v = m_mu[plane.h_v_indices]
    h_u = np.where(mu_u > 0, -e_v * (scale / np.where(mu_u > 0, mu_u, 1.0)), 0.0)
    h_v = np.where(mu_v > 0, +e_u * (scale / np.where(mu_v > 0, mu_v, 1.0)), 0.0)
    return h_u, h_v


def _dual_spacings(nodes: np.ndarray) -> np.ndarray:
    n = nodes.size
    if n < 2:
        return np.zeros(n)
    d = np.empty(n)
    d[1:-1] = 0.5 * (nodes[2:] - nodes[:-2])
    d[0] = 0.5 * (nodes[1] - nodes[0])
    d[-1] = 0.5 * (nodes[-1] - nodes[-2])
    return d


def _patch_duals(plane: PortPlane):
    u_nodes = np.unique(plane.v_edge_uv[:, 0])
    v_nodes = np.unique(plane.u_edge_uv[:, 1])
    dv = _dual_spacings(v_nodes)[np.searchsorted(v_nodes, plane.u_edge_uv[:, 1])]
    du = _dual_spacings(u_nodes)[np.searchsorted(u_nodes, plane.v_edge_uv[:, 0])]
    return dv, du


@dataclass
class PortModel:
    """A built modal port: plane, calibrated modes, scales and chain params."""

    name: str
    plane: PortPlane
    modes: list[DiscreteMode]
    dt: float
    omega_calc: float
    me_u_port: np.ndarray
    me_v_port: np.ndarray
    me_u_int: np.ndarray
    me_v_int: np.ndarray
    mh_u: np.ndarray
    mh_v: np.ndarray
    record_scale: np.ndarray
    source_scale: np.ndarray
    state_scale: np.ndarray
    dtbc_r: list
    dtbc_q: list
    dtbc_z0: list
    pair_spread: list
    mur_r: np.ndarray
    tau: np.ndarray
    slab_defect: float
    excitation: tuple | None = None  # (mode index, waveform callable)
    meta: dict = field(default_factory=dict)
    # complement absorber (factory.py:687-735): per plane edge [u | v] the
    # Mur coefficient r and the 0/1 live mask, or None
    comp_r: np.ndarray | None = None
    comp_live: np.ndarray | None = None

    @property
    def n_modes(self) -> int:
        return len(self.modes)

    @property
    def face(self) -> str:
        return self.plane.face

    @property
    def termination_kinds(self) -> list[str]:
        return ["dtbc" if r is not None else "mur" for r in self.dtbc_r]

    @property
    def complement_active(self) -> bool:
        """The absorber is configured and some channel runs on Mur
        (operator.py:1337-1346); fully certified ports keep the exact path."""
        return self.comp_r is not None and any(r is None for r in self.dtbc_r)

    @property
    def dtbc_line_params(self) -> dict:
        return {m: (r, self.dtbc_q[m], self.dtbc_z0[m]) for m, r in enumerate(self.dtbc_r)
                if r is not None}

    def set_excitation(self, mode_idx: int, waveform) -> None:
        if not (0 <= mode_idx < self.n_modes):
            raise ValueError(f"mode_idx {mode_idx} out of range [0, {self.n_modes})")
        self.excitation = (mode_idx, waveform)

    def clear_excitation(self) -> None:
        self.excitation = None

    def prof_e(self) -> np.ndarray:
        """(M, n_u + n_v) E profiles in [u | v] order."""
        return np.stack([np.concatenate([m.e_u, m.e_v]) for m in self.modes])

    def prof_h(self) -> np.ndarray:
        """(M, n_u + n_v) H profiles co-located with [u | v] (h_v | h_u)."""
        return np.stack([np.concatenate([m.h_v, m.h_u]) for m in self.modes])

    def packed_dtbc_kernels(self, n_steps: int, init: int = N_KERNEL_INIT) -> np.ndarray:
        rows = []
        for r, q in zip(self.dtbc_r, self.dtbc_q):
            rows.append(packed_kernels(r, q, n_steps, init) if r is not None else None)
        length = max((a.size for a in rows if a is not None), default=0)
        out = np.zeros((len(rows), length))
        for i, a in enumerate(rows):
            if a is not None:
                out[i, : a.size] = a
        return out


def _calibrate(plane, modes, me_u, me_v, mh_u, mh_v, omega_calc):
    """Numerical-path V/I calibration and physical sqrt(W) scales."""
    u_l, v_l, ndx = plane.u_edge_lengths, plane.v_edge_lengths, plane.normal_dx
    dv, du = _patch_duals(plane)
    area_u = u_l * dv * 1.0
    area_v = v_l * du * 1.0
    M = len(modes)
    rec = np.ones(M)
    src = np.ones(M)
    st = np.ones(M)
    out = []
    for m_idx, dm in enumerate(modes):
        v_test = float(np.dot(me_u, dm.e_u**2)) + float(np.dot(me_v, dm.e_v**2))
        i_test = float(np.dot(mh_u, dm.h_u**2)) + float(np.dot(mh_v, dm.h_v**2))
        Z = dm.z_modal(omega_calc)
        zr = float(Z.real)
        h_v_phys = dm.h_v * mh_v / (MU0 * ndx)
        h_u_phys = dm.h_u * mh_u / (MU0 * ndx)
        p_one = float(np.dot(dm.e_u * h_v_phys, area_u)) - float(np.dot(dm.e_v * h_u_phys, area_v))
        calibrated = False
        gamma = 0.0
        if i_test != 0.0 and abs(Z.imag) <= abs(Z.real) and zr != 0.0:
            gamma = (v_test / i_test) / zr
            if math.isfinite(gamma) and gamma != 0.0:
                out.append(replace(dm, h_u=dm.h_u * gamma, h_v=dm.h_v * gamma))
                calibrated = True
        if calibrated and p_one != 0.0 and math.isfinite(p_one):
            kappa = math.sqrt(abs(p_one) * abs(zr))
            H_v = dm.h_v * mh_v / (MU0 * ndx * u_l)
            H_u = dm.h_u * mh_u / (MU0 * ndx * v_l)
            s_h = float(np.dot(dm.e_u * H_v, area_u / u_l)) - float(np.dot(dm.e_v * H_u, area_v / v_l))
            hq = gamma * MU0 * ndx * (float(np.dot(dm.h_v * H_v, u_l)) + float(np.dot(dm.h_u * H_u, v_l)))
            c_state = 0.0
            if hq != 0.0:
                ratio = s_h / hq
                if math.isfinite(ratio) and ratio > 0.0:
                    c_state = math.sqrt(ratio) / kappa
            if c_state > 0.0 and math.isfinite(c_state):
                st[m_idx] = c_state
                rec[m_idx] = kappa * c_state
                src[m_idx] = math.sqrt(abs(zr)) / kappa / c_state
            else:
                rec[m_idx] = kappa
                src[m_idx] = math.sqrt(abs(zr)) / kappa
        if not calibrated:
            out.append(dm)
    return out, rec, src, st


def _chain_params(dm: DiscreteMode, me_u, me_v, mh_u, mh_v, dt, slab_defect):
    q_mode = dm.omega_c * dt
    if slab_defect is not None and slab_defect > DTBC_SLAB_DEFECT_TOL:
        return None, None, None
    pair = np.concatenate([me_u * mh_v, me_v * mh_u])
    weight = np.concatenate([me_u * dm.e_u**2, me_v * dm.e_v**2])
    w_max = float(weight.max()) if weight.size else 0.0
    active = weight > DTBC_WEIGHT_FLOOR * w_max
    if w_max <= 0.0 or not np.any(active) or np.any(pair[active] <= 0.0):
        return None, None, None
    r_pairs = dt / np.sqrt(pair[active])
    w = weight[active]
    r_mean = float(np.dot(w, r_pairs) / w.sum())
    spread = float(math.sqrt(np.dot(w, (r_pairs - r_mean) ** 2) / w.sum()) / r_mean)
    if spread > DTBC_PAIR_SPREAD_TOL or not (0.0 < r_mean <= 1.0):
        return None, None, spread
    return r_mean, q_mode, spread


def _chain_z0(dm: DiscreteMode, r_mode, me_u, me_v, mh_u, mh_v, dt) -> float:
    num = np.concatenate([mh_v * dm.h_v, mh_u * dm.h_u])
    den = np.concatenate([dm.e_u, dm.e_v])
    weight = np.concatenate([me_u * dm.e_u**2, me_v * dm.e_v**2])
    active = weight > DTBC_WEIGHT_FLOOR * float(weight.max())
    c_pairs = np.abs(num[active] / den[active])
    w = weight[active]
    c_pair = float(np.dot(w, c_pairs) / w.sum())
    n_v = float(np.dot(me_u, dm.e_u**2) + np.dot(me_v, dm.e_v**2))
    n_i = float(np.dot(mh_u, dm.h_u**2) + np.dot(mh_v, dm.h_v**2))
    return r_mode * n_v * c_pair / (dt * n_i)


def chain_slab_defect(dims, m_eps, m_mu, face: str) -> float:
    """Max relative slab deviation of the masses over the first feed cells."""
    nx, ny, nz = dims
    e_sh = [_e_shape(c, nx, ny, nz) for c in range(3)]
    h_sh = [_h_shape(c, nx, ny, nz) for c in range(3)]
    eo, ho = _offsets(e_sh), _offsets(h_sh)
    comps = [m_eps[eo[c]: eo[c] + int(np.prod(e_sh[c]))].reshape(e_sh[c]) for c in range(3)]
    comps += [m_mu[ho[c]: ho[c] + int(np.prod(h_sh[c]))].reshape(h_sh[c]) for c in range(3)]
    axis, is_max = FACES[face][0], FACES[face][1]
    defect = 0.0
    for arr in comps:
        n_along = arr.shape[axis]
        idx = range(n_along - 1, max(n_along - 4, -1), -1) if is_max else range(0, min(3, n_along))
        slabs = [np.take(arr, i, axis=axis) for i in idx]
        for a, b in zip(slabs, slabs[1:]):
            mask = (a != 0.0) | (b != 0.0)
            if not mask.any():
                continue
            den = np.maximum(np.abs(a[mask]), np.abs(b[mask]))
            defect = max(defect, float(np.max(np.abs(a[mask] - b[mask]) / den)))
    return defect


def warn_degenerate_cutoffs(modes, name: str, rtol: float = DEGENERATE_CUTOFF_RTOL) -> None:
    """UserWarning for two channels whose cut-offs agree within rtol: their
    flux inner product is not orthogonal, so the port couples them
    (factory.py:1318-1364).  TE_mn / TM_mn of a homogeneously filled PEC
    rectangle are such a pair at any resolution."""
    import warnings  # noqa: PLC0415

    ws = sorted(modes, key=lambda m: m.omega_c)
    for a, b in zip(ws, ws[1:]):
        if a.omega_c <= 0.0 or b.omega_c <= 0.0:
            continue
        rel = abs(b.omega_c - a.omega_c) / b.omega_c
        if rel < rtol:
            warnings.warn(f"modal port {name!r}: modes {a.name!r} (fc = {a.omega_c / 2e9 / math.pi:.4f} "
                          f"GHz) and {b.name!r} (fc = {b.omega_c / 2e9 / math.pi:.4f} GHz) are "
                          f"degenerate within {rel * 100:.3f}% relative cut-off difference; the "
                          "port couples them", UserWarning, stacklevel=3)


def build_port(name: str, face: str, mesh, m_eps: np.ndarray, m_mu: np.ndarray, *, dt: float,
               f_calc: float, n_modes: int = 1, eps_r: float = 1.0,
               termination: str = "auto") -> PortModel:
    """Numerical modal port (TE and TM) on a whole bbox face (factory.py:1367-1877 order)."""
    plane, m_eps, m_mu, pec = _port_inputs(face, mesh, m_eps, m_mu, dt, f_calc)
    modes = solve_modes(plane, mesh.grid, m_eps, m_mu, pec, n_modes, f_calc, eps_r)
    warn_degenerate_cutoffs(modes, name)
    return _assemble_port(name, plane, mesh, m_eps, m_mu, modes, dt, f_calc, termination)


def _port_inputs(face, mesh, m_eps, m_mu, dt, f_calc):
    """Feed check, plane, and the masses and PEC mask flattened at the port
    face (the builder-local flatten of factory.py:1419-1458)."""
    if dt <= 0.0 or f_calc <= 0.0:
        raise ValueError("dt and f_calc must be positive")
    validate_feed(face, mesh.grid)
    plane = PortPlane.from_mesh(face, mesh.grid)
    dims = mesh.dims
    m_eps = _mesh.flatten_port_mass(m_eps, dims, face)
    m_mu = _mesh.flatten_port_mu(m_mu, dims, face)
    pec = _mesh.pec_flat(_mesh.flatten_port_pec(mesh.pec_mask_edges, dims, face), dims)
    return plane, m_eps, m_mu, pec


def _assemble_port(name, plane, mesh, m_eps, m_mu, modes, dt, f_calc, termination) -> PortModel:
    """Calibration, Mur and DTBC chain parameters and the complement absorber
    of a port whose modes are solved (factory.py:1803-1877, operator.py:596-941)."""
    omega_calc = 2.0 * math.pi * f_calc
    me_u, me_v = m_eps[plane.e_u_indices], m_eps[plane.e_v_indices]
    me_ui, me_vi = m_eps[plane.e_u_indices_interior], m_eps[plane.e_v_indices_interior]
    mh_u, mh_v = m_mu[plane.h_u_indices], m_mu[plane.h_v_indices]
    defect = chain_slab_defect(mesh.dims, m_eps, m_mu, plane.face)
    modes, rec, src, st = _calibrate(plane, modes, me_u, me_v, mh_u, mh_v, omega_calc)
    M = len(modes)
    v_p = np.empty(M)
    mur_r = np.empty(M)
    tau = np.empty(M)
    for m, dm in enumerate(modes):
        beta = abs(dm.gamma(omega_calc).imag)
        v_p[m] = (omega_calc / beta) if beta > 0.0 else C0
        mur_r[m] = (v_p[m] * dt - plane.normal_dx) / (v_p[m] * dt + plane.normal_dx)
        tau[m] = plane.normal_dx / v_p[m]
    rs, qs, z0s, spreads = [], [], [], []
    if termination not in ("auto", "mur"):
        raise ValueError(f"termination must be 'auto' or 'mur', got {termination!r}")
    for dm in modes:
        r_m, q_m, spread = _chain_params(dm, me_u, me_v, mh_u, mh_v, dt, defect)
        use = termination == "auto" and r_m is not None
        rs.append(r_m if use else None)
        qs.append(q_m if use else None)
        z0s.append(_chain_z0(dm, r_m, me_u, me_v, mh_u, mh_v, dt) if use and q_m > 0.0 else None)
        spreads.append(spread)
    comp_r, comp_live = complement_absorber(plane, mesh, m_eps, dt)
    return PortModel(name=name, plane=plane, modes=modes, dt=dt, omega_calc=omega_calc,
                     me_u_port=me_u, me_v_port=me_v, me_u_int=me_ui, me_v_int=me_vi, mh_u=mh_u,
                     mh_v=mh_v, record_scale=rec, source_scale=src, state_scale=st, dtbc_r=rs,
                     dtbc_q=qs, dtbc_z0=z0s, pair_spread=spreads, mur_r=mur_r, tau=tau,
                     slab_defect=defect, comp_r=comp_r, comp_live=comp_live)


# ---------------------------------------------------------------------------
# TEM ports of multi-conductor feeds (coax)

def _components(a: np.ndarray, b: np.ndarray, n: int) -> dict:
    """Connected components of the node links a-b: label -> sorted nodes
    (nodes without a link are in none)."""
    from scipy.sparse.csgraph import connected_components  # noqa: PLC0415

    if a.size == 0:
        return {}
    rows, cols = np.concatenate([a, b]), np.concatenate([b, a])
    adj = sp.csr_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(n, n))
    _, labels = connected_components(adj, directed=False)
    out: dict = {}
    for node in np.unique(rows):
        out.setdefault(int(labels[node]), []).append(int(node))
    return out


def conductor_groups(plane: PortPlane, mesh, pec_e_flat: np.ndarray) -> list[np.ndarray]:
    """[ground, signal_1, ...] plane-node groups of the conductors on a port
    plane (auto_conductors.py:54-247): connected components of the nodes
    joined by PEC plane edges; PEC cells of the port-adjacent slab link their
    four corners, which only fuses edge components into one conductor (or,
    with fewer than two edge components, supplies the node sets, with a
    warning).  Ground is the largest component touching the plane's
    boundary (else the largest), the rest by descending size."""
    import warnings  # noqa: PLC0415

    G = gradient_2d(plane)
    nu1, nv1 = plane.u_node_window[1] + 1, plane.v_node_window[1] + 1
    prim = np.concatenate([plane.e_u_indices, plane.e_v_indices])
    rows = G[np.asarray(pec_e_flat, bool)[prim]].tocoo()
    order = np.lexsort((rows.data, rows.row))  # per edge: start (-1) then end (+1)
    cols = rows.col[order].reshape(-1, 2)
    ea, eb = cols[:, 0].astype(np.int64), cols[:, 1].astype(np.int64)
    comps = _components(ea, eb, nu1 * nv1)
    # PEC cells of the slab next to the plane link their corners
    n_ax, is_max, u_ax, v_ax = FACES[plane.face]
    slab = np.take(mesh.material_id, mesh.dims[n_ax] - 1 if is_max else 0, axis=n_ax)
    if u_ax > v_ax:
        slab = slab.T
    pec_ids = [k for k, m in mesh.material_library.items() if getattr(m, "is_pec", False)]
    cu, cv = np.nonzero(np.isin(slab, pec_ids)) if pec_ids else (np.zeros(0, int), np.zeros(0, int))
    if cu.size:
        c00 = cu * nv1 + cv
        ca = np.concatenate([c00, c00, c00])
        cb = np.concatenate([c00 + nv1, c00 + 1, c00 + nv1 + 1])
        merged = _components(np.concatenate([ea, ca]), np.concatenate([eb, cb]), nu1 * nv1)
        if len(comps) >= 2:
            lab = {n: l for l, nodes in merged.items() for n in nodes}
            fused: dict = {}
            for nodes in comps.values():
                fused.setdefault(lab[nodes[0]], []).extend(nodes)
            comps = {k: sorted(v) for k, v in fused.items()}
        else:
            comps = merged
            if len(merged) >= 2:
                warnings.warn(f"port {plane.face}: conductors resolved from PEC cells only; refine the "
                              "transversal mesh", UserWarning, stacklevel=2)
    if len(comps) < 2:
        raise ValueError(f"port {plane.face}: {len(comps)} PEC-connected conductor(s) on the plane; "
                         "a TEM port needs a ground and at least one signal conductor")

    def touches(nodes):
        iu, iv = np.divmod(np.asarray(nodes), nv1)
        return bool(np.any((iu == 0) | (iu == nu1 - 1) | (iv == 0) | (iv == nv1 - 1)))

    wall = [k for k, v in comps.items() if touches(v)]
    ground = max(wall or list(comps), key=lambda k: len(comps[k]))
    rest = sorted((k for k in comps if k != ground), key=lambda k: -len(comps[k]))
    return [np.asarray(comps[ground], np.int64)] + [np.asarray(comps[k], np.int64) for k in rest]


def _boundary_window_factor(plane: PortPlane, grid, pmc_faces=()) -> np.ndarray:
    """Per plane edge [u | v]: the share of its dual cell the plane's
    capacitance owns (tem_laplace.py:170-306): 1 inside, 1/2 on the plane's
    boundary lines (a wall on the line), 1 on a PMC face's line (the
    magnetic wall sits half a cell beyond it)."""
    n_ax, is_max, u_ax, v_ax = FACES[plane.face]
    names = (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax"))
    nu, nv = plane.u_node_window[1], plane.v_node_window[1]

    def f(axis, lo):
        return 1.0 if names[axis][0 if lo else 1] in pmc_faces else 0.5

    fu = np.ones((nu, nv + 1))
    fu[:, 0], fu[:, -1] = f(v_ax, True), f(v_ax, False)
    fv = np.ones((nv, nu + 1))
    fv[:, 0], fv[:, -1] = f(u_ax, True), f(u_ax, False)
    return np.concatenate([fu.ravel(), fv.ravel()])


def solve_tem_modes(plane: PortPlane, grid, m_eps, m_mu, groups, eps_r: float = 1.0,
                    pmc_faces=()) -> list[DiscreteMode]:
    """TEM channels of a multi-conductor plane by the FIT Laplace with
    conductor potentials (tem_laplace.py:321-488, :640-771): per signal
    conductor k, phi = 1 on k, 0 on the others, L phi = 0 elsewhere with
    L = G^T M_eps G; e = -G phi, M_eps-normalised, signed so the D flux out
    of conductor k is positive; C' = e_raw^T M_cap e_raw / dx_n with the
    boundary-corrected mass, Z_line = sqrt(eps_r) / (c C').  Several signal
    conductors: the eigenvectors of the raw fields' Gram matrix (descending,
    the leading near-largest conductor weight positive)."""
    from scipy.sparse.linalg import spsolve  # noqa: PLC0415

    if eps_r <= 0.0:
        raise ValueError("epsilon_r must be positive")
    if len(groups) < 2:
        raise ValueError("a TEM port needs a ground and at least one signal conductor")
    G = gradient_2d(plane)
    prim = np.concatenate([plane.e_u_indices, plane.e_v_indices])
    Md = np.asarray(m_eps[prim], float)
    Mcap = Md * _boundary_window_factor(plane, grid, pmc_faces)
    n_nodes = G.shape[1]
    fixed = np.concatenate(groups)
    if np.unique(fixed).size != fixed.size:
        raise ValueError("conductor groups must be disjoint")
    free = np.setdiff1d(np.arange(n_nodes), fixed)
    if free.size == 0:
        raise RuntimeError("every plane node is on a conductor")
    L = (G.T @ sp.diags(Md) @ G).tocsc()
    L_ff = L[free][:, free]
    L_fb = L[free][:, fixed]
    owner = np.concatenate([np.full(g.size, k) for k, g in enumerate(groups)])
    raw = []
    for k in range(1, len(groups)):
        phi_b = (owner == k).astype(float)
        phi = np.zeros(n_nodes)
        phi[free] = spsolve(L_ff, -(L_fb @ phi_b))
        phi[fixed] = phi_b
        e_raw = -(G @ phi)
        norm_sq = float(e_raw @ (Md * e_raw))
        if norm_sq <= 0.0:
            raise RuntimeError(f"signal conductor {k}: degenerate TEM field")
        e_n = e_raw / math.sqrt(norm_sq)
        outward = -float((G.T @ (Md * e_n))[groups[k]].sum())
        if outward < 0.0:
            e_n, e_raw = -e_n, -e_raw
        raw.append((e_n, float(e_raw @ (Mcap * e_raw)), e_raw))
    ndx = float(plane.normal_dx)
    if len(raw) == 1:
        channels = [(raw[0][0], raw[0][1] / ndx)]
    else:
        E = np.stack([r[2] for r in raw], axis=1)
        g_fit = E.T @ (Md[:, None] * E)
        g_cap = E.T @ (Mcap[:, None] * E)
        w_all, W = np.linalg.eigh(g_fit)
        order = np.argsort(w_all)[::-1]
        w_all, W = w_all[order], W[:, order]
        channels = []
        for m in range(W.shape[1]):
            w = W[:, m]
            aw = np.abs(w)
            lead = int(np.flatnonzero(aw >= (1.0 - 1e-9) * aw.max())[0])
            if w[lead] < 0.0:
                w = -w
            channels.append(((E @ w) / math.sqrt(float(w_all[m])), float(w @ g_cap @ w) / ndx))
    n_u = plane.e_u_indices.size
    z_tem = ETA0 / math.sqrt(eps_r)
    out = []
    for i, (e, c_pl) in enumerate(channels):
        e_u, e_v = e[:n_u].copy(), e[n_u:].copy()
        h_u, h_v = travelling_wave_h(e_u, e_v, plane, m_mu, z_tem)
        out.append(DiscreteMode(f"TEM_lap{i:02d}", 0.0, eps_r, e_u, e_v, h_u, h_v, "TEM",
                                z_line=math.sqrt(eps_r) / (C0 * c_pl)))
    return out


def conductor_centroid(plane: PortPlane, grid, nodes) -> tuple[float, float]:
    """(u, v) mean of a conductor's plane-node coordinates (factory.py:953-971)."""
    _, _, u_ax, v_ax = FACES[plane.face]
    nv1 = plane.v_node_window[1] + 1
    coords = (grid.x, grid.y, grid.z)
    iu, iv = np.divmod(np.asarray(nodes), nv1)
    return float(np.mean(coords[u_ax][iu])), float(np.mean(coords[v_ax][iv]))


def build_coax_port(name: str, face: str, mesh, m_eps: np.ndarray, m_mu: np.ndarray, *, dt: float,
                    f_calc: float, center, eps_r: float = 1.0, n_modes: int = 1,
                    termination: str = "auto", pmc_faces=()) -> PortModel:
    """Coaxial (TEM) modal port on a whole bbox face: the FIT path of the
    reference's coax port (factory.py:1470-1520) — conductors found on the
    flattened PEC mask, exactly two of them centred within one cell of
    ``center`` (global tangential frame, lower axis first; :974-1020), the
    Laplace TEM channel, then the same calibration, chain and absorber set-up
    as every modal port."""
    plane, m_eps, m_mu, pec = _port_inputs(face, mesh, m_eps, m_mu, dt, f_calc)
    groups = conductor_groups(plane, mesh, pec)
    if len(groups) != 2:
        raise ValueError(f"coax port {name!r} on {face!r}: {len(groups)} conductor groups found; "
                         "expected 2 (inner and outer)")
    _, _, u_ax, v_ax = FACES[face]
    cu, cv = (center[0], center[1]) if u_ax < v_ax else (center[1], center[0])
    d = (mesh.grid.dx, mesh.grid.dy, mesh.grid.dz)
    tol = max(float(np.max(d[u_ax])), float(np.max(d[v_ax])))
    for k, g in enumerate(groups):
        gu, gv = conductor_centroid(plane, mesh.grid, g)
        if math.hypot(gu - cu, gv - cv) > tol:
            raise ValueError(f"coax port {name!r}: {'outer' if k == 0 else 'inner'} conductor centroid "
                             f"({gu:.6g}, {gv:.6g}) is more than one cell from the centre ({cu:.6g}, {cv:.6g})")
    modes = solve_tem_modes(plane, mesh.grid, m_eps, m_mu, groups, eps_r, pmc_faces)[:n_modes]
    return _assemble_port(name, plane, mesh, m_eps, m_mu, modes, dt, f_calc, termination)


def complement_absorber(plane: PortPlane, mesh, m_eps_flat: np.ndarray, dt: float):
    """Per plane edge [u | v]: the Mur coefficient of the complement absorber
    r = (c dt - dx_n)/(c dt + dx_n) with the local phase velocity
    c = C0 / sqrt(M_eps / M_eps_vacuum) of the port-flattened masses (0 where
    M_eps <= 0), and the live mask (neither PEC in the mesh's own edge mask
    nor frozen) (factory.py:687-735)."""
    from .dispersion import _geom_e  # noqa: PLC0415

    dims = mesh.dims
    m_vac = _mesh.flatten_port_mass(_mesh.EPS0 * _geom_e(mesh.grid), dims, plane.face)
    idx = np.concatenate([plane.e_u_indices, plane.e_v_indices])
    m_e = np.asarray(m_eps_flat[idx], float)
    dead = m_e <= 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        c_loc = C0 / np.sqrt(m_e / np.asarray(m_vac[idx], float))
        r = (c_loc * dt - plane.normal_dx) / (c_loc * dt + plane.normal_dx)
    r = np.where(dead, 0.0, r)
    pec = _mesh.pec_flat(mesh.pec_mask_edges, dims)[idx] | dead
    return r, (~pec).astype(float)


def build_port_local(name: str, face: str, mesh, *, dt: float, f_calc: float, n_modes: int = 1,
                     eps_r: float = 1.0, termination: str = "auto", feed_cells: int = 8) -> PortModel:
    """``build_port`` evaluated on the feed slab only.

    Everything a port model holds — plane masses, PEC mask, modes,
    calibration, chain parameters, the slab defect of the first feed cells —
    depends on the ``feed_cells`` cell layers next to the face alone: each
    mass entry on those layers comes from its owning cell and the local
    widths.  The model is built on a sub-mesh of those layers (same grid
    lines, material ids and walls; the cut face is left open) and its plane
    is re-targeted to the full grid's flat indices, so it equals the
    full-grid build bit for bit without the full-grid mass matrices.
    A mesh whose PEC mask was edited by hand uses the full-grid build.
    """
    axis, is_max = FACES[face][0], FACES[face][1]
    N = mesh.dims[axis]
    if feed_cells < 4:
        raise ValueError("feed_cells must be at least 4")
    if feed_cells >= N or not getattr(mesh, "pec_rule", False):
        return build_port(name, face, mesh, _mesh.build_m_eps(mesh), _mesh.build_m_mu(mesh), dt=dt,
                          f_calc=f_calc, n_modes=n_modes, eps_r=eps_r, termination=termination)
    g = mesh.grid
    lines = [g.x, g.y, g.z]
    lines[axis] = lines[axis][N - feed_cells:] if is_max else lines[axis][: feed_cells + 1]
    sub_grid = _mesh.GridLines(*lines)
    sl = [slice(None)] * 3
    sl[axis] = slice(N - feed_cells, N) if is_max else slice(0, feed_cells)
    mid = np.ascontiguousarray(mesh.material_id[tuple(sl)])
    cut = ("xmin", "ymin", "zmin")[axis] if is_max else ("xmax", "ymax", "zmax")[axis]
    walls = tuple(f for f in mesh.pec_faces if f != cut)
    pec = _mesh.build_pec_mask(sub_grid, mid, mesh.material_library)
    _mesh.or_bbox_pec_walls(pec, *sub_grid.dims, walls)
    sub = _mesh.Mesh(grid=sub_grid, material_id=mid, material_library=mesh.material_library,
                     pec_mask_edges=pec, pec_faces=walls)
    P = build_port(name, face, sub, _mesh.build_m_eps(sub), _mesh.build_m_mu(sub), dt=dt,
                   f_calc=f_calc, n_modes=n_modes, eps_r=eps_r, termination=termination)
    full = PortPlane.from_mesh(face, g)
    local = P.plane
    if not (np.array_equal(full.u_edge_uv, local.u_edge_uv) and np.array_equal(full.v_edge_uv, local.v_edge_uv)
            and full.coordinate == local.coordinate and full.normal_dx == local.normal_dx):
        raise RuntimeError(f"port {name}: feed sub-mesh plane does not match the full grid")
    P.plane = full
    return P
