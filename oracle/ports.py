"""Oracle: modal port plane, discrete TE modes, calibration and port step.

Restates (paths relative to the reference repository root):

* plane maps        src/magnelio/ports/_modal/port_plane.py:32-334 and the
                    face frames of src/magnelio/mesh/faces.py:63-71
* TEM channels     taken as given (profiles and line impedance); the H
                    profiles use the wave impedance eta0/sqrt(eps_r), the
                    calibration the line impedance, the chain q = 0 with no
                    static z0 (ports/_modal/tem_laplace.py:474-482,
                    operator.py:666-668, :695-697)
* discrete TM modes closed form on a uniform PEC rectangle: the node
                    potential phi = sin(m pi ia/Na) sin(n pi ib/Nb) of the
                    normal-E problem (ports/_modal/curl_curl_2d.py:127-218),
                    transverse profile its plane gradient e_t = G phi
                    (numerical_2d.py:585-753), w_c^2 = c^2 (km^2 + kn^2)
                    with the km, kn below; TM wave impedance
                    gamma/(j w eps0 eps_r) and |Z| bake-in
                    sqrt|kc^2 - k^2|/(w eps0 eps_r) (mode.py:206-211,
                    numerical_2d.py:828-853)
* discrete TE modes closed form on a uniform PEC rectangle: edge voltages
                    e_x = kn cos(m pi (i+1/2)/Nx) sin(n pi j/Ny) dx,
                    e_y = -km sin(m pi i/Nx) cos(n pi (j+1/2)/Ny) dy,
                    km = (2/dx) sin(m pi/(2Nx)), kn = (2/dy) sin(n pi/(2Ny)),
                    w_c^2 = c^2 (km^2 + kn^2) — the exact eigenpairs of the
                    2-D curl-curl pencil the reference solves with eigsh
                    (ports/_modal/numerical_2d.py:454-583), M_eps-normalised,
                    sign by the port-symmetric rule :758-802
* H profiles        ports/_modal/tem_laplace.py:102-167
* calibration       ports/_modal/operator.py:778-941 (numerical path)
* chain r, q, z0    ports/_modal/operator.py:671-776
* port step         ports/_modal/operator.py:1351-1536 (Mur, excited Mur
                    TF/SF with the ring-buffer look-back :1539-1562, DTBC)
* projections       ports/_modal/operator.py:1277-1321
* recorder          src/magnelio/ports/recorder.py:216-250
* complement        ports/_modal/factory.py:687-735 (per-edge Mur coefficient
                    r = (c dt - dx_n)/(c dt + dx_n), c = C0/sqrt(M_eps/M_eps_vac)
                    on the plane edges of the flattened masses, 0 where
                    M_eps <= 0; live = not PEC (the mesh's own mask) and not
                    frozen), operator.py:1395-1427, :1499-1522 (live while a
                    channel of the port runs on Mur: the interior-plane
                    remainder comp = e_int - sum_m V_int,m p_m, advanced by
                    live (comp_int_prev + r (comp - comp_port_prev)) and added
                    to the modal write)

Parity pins (tests/test_oracle_ports.py): the closed-form modes are
eigenpairs of an independently assembled curl-curl pencil to 1e-12, are
M_eps-orthonormal and divergence-free; the calibration makes the unit
travelling wave satisfy V/I = Z_modal exactly; the discrete TE10 chain
parameters satisfy r = c dt / dz and q = w_c dt.  The closed-form TM modes
are eigenpairs of the normal-E pencil assembled from the oracle's own 3-D
curl operators, curl-free on the plane and M_eps-orthogonal to every TE
mode.  ``record_scale`` /
``source_scale`` / ``state_scale`` have no closed form; they are pinned by
the reference's own values on the uniform WR-90 guide
(tests/test_golden_ports.py, tests/golden/ports_wr90.npz), and the whole
port step by the reference's 300-step runs (tests/test_golden_c1.py).
"""

from __future__ import annotations

import collections
import math

import numpy as np

from .dtbc import DTBC

C0 = 299_792_458.0
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0 * C0)

FRAMES = {"xmin": (0, False, 1, 2), "xmax": (0, True, 2, 1), "ymin": (1, False, 2, 0),
          "ymax": (1, True, 0, 2), "zmin": (2, False, 0, 1), "zmax": (2, True, 1, 0)}


def _eshape(c, n):
    nx, ny, nz = n
    return [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)][c]


def _hshape(c, n):
    nx, ny, nz = n
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)][c]


def plane(face: str, x, y, z) -> dict:
    """Index maps and geometry of a whole-face port plane (reference flat layout)."""
    n_ax, is_max, u_ax, v_ax = FRAMES[face]
    nodes = (np.asarray(x), np.asarray(y), np.asarray(z))
    N = tuple(a.size - 1 for a in nodes)
    d = tuple(np.diff(a) for a in nodes)
    cen = tuple(0.5 * (a[:-1] + a[1:]) for a in nodes)
    e_off = np.cumsum([0] + [int(np.prod(_eshape(c, N))) for c in range(3)])
    h_off = np.cumsum([0] + [int(np.prod(_hshape(c, N))) for c in range(3)])
    k_e = N[n_ax] if is_max else 0
    k_i = N[n_ax] - 1 if is_max else 1
    k_h = N[n_ax] - 1 if is_max else 0

    def flat(shape, off, ijk):
        return off + (ijk[0] * shape[1] + ijk[1]) * shape[2] + ijk[2]

    def fam(p_ax, s_ax, kn):
        e_idx, h_idx, uv, ln = [], [], [], []
        for a in range(N[p_ax]):  # primal cells
            for b in range(N[s_ax] + 1):  # secondary nodes
                ijk = [0, 0, 0]
                ijk[n_ax], ijk[p_ax], ijk[s_ax] = kn, a, b
                e_idx.append(flat(_eshape(p_ax, N), e_off[p_ax], ijk))
                ijk[n_ax] = k_h
                h_idx.append(flat(_hshape(s_ax, N), h_off[s_ax], ijk))
                uv.append((cen[p_ax][a], nodes[s_ax][b]))
                ln.append(d[p_ax][a])
        return (np.array(e_idx, np.int64), np.array(h_idx, np.int64), np.array(uv), np.array(ln))

    eu, hv, uuv, ul = fam(u_ax, v_ax, k_e)
    eui, _, _, _ = fam(u_ax, v_ax, k_i)
    ev, hu, vuv, vl = fam(v_ax, u_ax, k_e)
    evi, _, _, _ = fam(v_ax, u_ax, k_i)
    return dict(face=face, e_u=eu, h_v=hv, u_uv=uuv, u_len=ul, e_v=ev, h_u=hu, v_uv=vuv[:, ::-1],
                v_len=vl, e_u_int=eui, e_v_int=evi, normal_dx=float(d[n_ax][k_h]),
                u_axis=u_ax, v_axis=v_ax, n_axis=n_ax, N=N)


def resolve_sign(pl: dict, v: np.ndarray) -> np.ndarray:
    if pl["u_axis"] < pl["v_axis"]:
        lo = np.concatenate([pl["u_uv"][:, 0], pl["v_uv"][:, 0]])
        hi = np.concatenate([pl["u_uv"][:, 1], pl["v_uv"][:, 1]])
    else:
        lo = np.concatenate([pl["u_uv"][:, 1], pl["v_uv"][:, 1]])
        hi = np.concatenate([pl["u_uv"][:, 0], pl["v_uv"][:, 0]])
    dlo = lo - 0.5 * (lo.max() + lo.min())
    dhi = hi - 0.5 * (hi.max() + hi.min())
    vmax = float(np.max(np.abs(v)))
    for f in (dlo, dhi, dlo * dhi, dlo * dlo + dhi * dhi):
        ov = float(np.sum(v * f))
        if abs(ov) > 1e-9 * vmax * float(np.max(np.abs(f))) * v.size:
            return -v if ov < 0 else v
    j = int(np.argmax(np.abs(v)))
    return -v if v[j] < 0 else v


def rect_te_mode(pl: dict, x, y, z, m: int, n: int, m_eps: np.ndarray, eps_r: float = 1.0,
                 window=None):
    """Closed-form discrete TE_mn on a uniform PEC-walled plane.

    m counts along the lower-numbered global tangential axis, n along the
    higher one.  ``window``: ((a0, a1), (b0, b1)) node-index ranges of a
    rectangular guide inside the plane (PEC around it; default the whole
    plane); the mode vanishes outside.  Returns (e_u, e_v, omega_c),
    M_eps-normalised and signed.
    """
    nodes = (np.asarray(x), np.asarray(y), np.asarray(z))
    ta = sorted([pl["u_axis"], pl["v_axis"]])  # (a, b) global tangential axes
    (ia0, ia1), (ib0, ib1) = window if window is not None else ((0, pl["N"][ta[0]]), (0, pl["N"][ta[1]]))
    Na, Nb = ia1 - ia0, ib1 - ib0
    da = float(np.diff(nodes[ta[0]])[0])
    db = float(np.diff(nodes[ta[1]])[0])
    a0, b0 = nodes[ta[0]][0], nodes[ta[1]][0]
    km = (2.0 / da) * math.sin(m * math.pi / (2 * Na))
    kn = (2.0 / db) * math.sin(n * math.pi / (2 * Nb))

    def prof(uv, along_a):
        # uv columns are (primal-axis midpoint, secondary node) of the family
        # the sine factor sits on wall nodes (window index 0 or N): exactly zero there
        if along_a:  # edges along axis a: E_a = kn cos(m pi ia/Na) sin(n pi ib/Nb)
            ia = (uv[:, 0] - a0) / da - ia0
            ib = np.rint((uv[:, 1] - b0) / db) - ib0
            s = np.where((ib <= 0) | (ib >= Nb) | (ia < 0) | (ia > Na), 0.0, np.sin(n * math.pi * ib / Nb))
            return kn * np.cos(m * math.pi * ia / Na) * s * da
        ib = (uv[:, 0] - b0) / db - ib0
        ia = np.rint((uv[:, 1] - a0) / da) - ia0
        s = np.where((ia <= 0) | (ia >= Na) | (ib < 0) | (ib > Nb), 0.0, np.sin(m * math.pi * ia / Na))
        return -km * s * np.cos(n * math.pi * ib / Nb) * db

    u_uv = pl["u_uv"]
    v_uv_raw = pl["v_uv"][:, ::-1]  # back to (primal, secondary) of the v family
    e_u = prof(u_uv, pl["u_axis"] == ta[0])
    e_v = prof(v_uv_raw, pl["v_axis"] == ta[0])
    full = np.concatenate([e_u, e_v])
    me = np.concatenate([m_eps[pl["e_u"]], m_eps[pl["e_v"]]])
    full = full / math.sqrt(float(np.sum(me * full * full)))
    full = resolve_sign(pl, full)
    nu = pl["e_u"].size
    wc = C0 / math.sqrt(eps_r) * math.sqrt(km * km + kn * kn)
    return full[:nu].copy(), full[nu:].copy(), wc


def rect_tm_mode(pl: dict, x, y, z, m: int, n: int, m_eps: np.ndarray, eps_r: float = 1.0):
    """Closed-form discrete TM_mn (m, n >= 1) on a uniform PEC-walled plane.

    The transverse profile is the plane gradient of the node potential
    phi(ia, ib) = sin(m pi ia/Na) sin(n pi ib/Nb): on an edge along axis a
    at (cell ia, node ib), phi(ia+1, ib) - phi(ia, ib) =
    km da cos(m pi (ia+1/2)/Na) sin(n pi ib/Nb), and likewise along b.
    Returns (e_u, e_v, omega_c), M_eps-normalised and signed.
    """
    if m < 1 or n < 1:
        raise ValueError("TM_mn needs m, n >= 1")
    nodes = (np.asarray(x), np.asarray(y), np.asarray(z))
    ta = sorted([pl["u_axis"], pl["v_axis"]])
    Na, Nb = pl["N"][ta[0]], pl["N"][ta[1]]
    da = float(np.diff(nodes[ta[0]])[0])
    db = float(np.diff(nodes[ta[1]])[0])
    a0, b0 = nodes[ta[0]][0], nodes[ta[1]][0]
    km = (2.0 / da) * math.sin(m * math.pi / (2 * Na))
    kn = (2.0 / db) * math.sin(n * math.pi / (2 * Nb))

    def prof(uv, along_a):
        if along_a:
            ia = (uv[:, 0] - a0) / da
            ib = np.rint((uv[:, 1] - b0) / db)
            s = np.where((ib == 0) | (ib == Nb), 0.0, np.sin(n * math.pi * ib / Nb))
            return km * da * np.cos(m * math.pi * ia / Na) * s
        ib = (uv[:, 0] - b0) / db
        ia = np.rint((uv[:, 1] - a0) / da)
        s = np.where((ia == 0) | (ia == Na), 0.0, np.sin(m * math.pi * ia / Na))
        return kn * db * s * np.cos(n * math.pi * ib / Nb)

    e_u = prof(pl["u_uv"], pl["u_axis"] == ta[0])
    e_v = prof(pl["v_uv"][:, ::-1], pl["v_axis"] == ta[0])
    full = np.concatenate([e_u, e_v])
    me = np.concatenate([m_eps[pl["e_u"]], m_eps[pl["e_v"]]])
    full = full / math.sqrt(float(np.sum(me * full * full)))
    full = resolve_sign(pl, full)
    nu = pl["e_u"].size
    wc = C0 / math.sqrt(eps_r) * math.sqrt(km * km + kn * kn)
    return full[:nu].copy(), full[nu:].copy(), wc


def z_te(omega, omega_c, eps_r=1.0):
    k2 = omega**2 * MU0 * EPS0 * eps_r
    kc2 = omega_c**2 * MU0 * EPS0 * eps_r
    diff = kc2 - k2
    g = 1j * math.sqrt(-diff) if diff <= 0 else complex(math.sqrt(diff))
    return 1j * omega * MU0 / g, g


def z_mode(kind, omega, omega_c, eps_r=1.0, z_line=None):
    """(Z, gamma) of a TE, TM or TEM channel (a TEM channel's Z is its line
    impedance, mode.py:215-231)."""
    z, g = z_te(omega, omega_c, eps_r)
    if kind == "TEM":
        return complex(z_line if z_line is not None else math.sqrt(MU0 / (EPS0 * eps_r))), g
    if kind == "TM":
        return g / (1j * omega * EPS0 * eps_r), g
    return z, g


def z_real(kind, omega, omega_c, eps_r=1.0):
    """|Z| at omega used to bake the H profiles (TEM: the wave impedance
    eta0 / sqrt(eps_r), tem_laplace.py:474-482)."""
    if kind == "TEM":
        return math.sqrt(MU0 / (EPS0 * eps_r))
    root = math.sqrt(abs(omega_c**2 * MU0 * EPS0 * eps_r - omega**2 * MU0 * EPS0 * eps_r))
    return omega * MU0 / root if kind == "TE" else root / (omega * EPS0 * eps_r)


class Port:
    """Oracle modal port: calibrated TE / TM channels, Mur/DTBC termination."""

    def __init__(self, name, pl, modes, m_eps, m_mu, dt, omega_calc, eps_r=1.0, termination="auto",
                 complement=None, kernel_init=None):
        """modes: list of (e_u, e_v, omega_c) from :func:`rect_te_mode`, or
        (e_u, e_v, omega_c, "TE" | "TM") (:func:`rect_tm_mode` is "TM");
        termination "auto" (exact DTBC) or "mur" (modal Mur-1 everywhere);
        complement: (r_u, r_v, live_u, live_v) of :func:`complement_arrays`
        or None; kernel_init: first DTBC kernel length (None: the
        reference's 4096)."""
        self.name, self.pl, self.dt = name, pl, dt
        self.comp = None if complement is None else tuple(np.asarray(a, float) for a in complement)
        self.comp_int_u = np.zeros(pl["e_u"].size)
        self.comp_int_v = np.zeros(pl["e_v"].size)
        self.comp_port_u = np.zeros(pl["e_u"].size)
        self.comp_port_v = np.zeros(pl["e_v"].size)
        self.me_u, self.me_v = m_eps[pl["e_u"]], m_eps[pl["e_v"]]
        self.me_ui, self.me_vi = m_eps[pl["e_u_int"]], m_eps[pl["e_v_int"]]
        self.mh_u, self.mh_v = m_mu[pl["h_u"]], m_mu[pl["h_v"]]
        ndx = pl["normal_dx"]
        self.modes = []
        M = len(modes)
        self.record_scale = np.ones(M)
        self.source_scale = np.ones(M)
        self.state_scale = np.ones(M)
        # patch duals for the physical-power surrogate
        u_nodes = np.unique(pl["v_uv"][:, 0])
        v_nodes = np.unique(pl["u_uv"][:, 1])

        def duals(nodes_):
            dd = np.empty(nodes_.size)
            dd[1:-1] = 0.5 * (nodes_[2:] - nodes_[:-2])
            dd[0] = 0.5 * (nodes_[1] - nodes_[0])
            dd[-1] = 0.5 * (nodes_[-1] - nodes_[-2])
            return dd

        dv = duals(v_nodes)[np.searchsorted(v_nodes, pl["u_uv"][:, 1])]
        du = duals(u_nodes)[np.searchsorted(u_nodes, pl["v_uv"][:, 0])]
        area_u, area_v = pl["u_len"] * dv * 1.0, pl["v_len"] * du * 1.0
        for mi, mode in enumerate(modes):
            e_u, e_v, wc = mode[:3]
            kind = mode[3] if len(mode) > 3 else "TE"
            z_line = mode[4] if len(mode) > 4 else None
            sc = MU0 * ndx / z_real(kind, omega_calc, wc, eps_r)
            h_u = np.where(self.mh_u > 0, -e_v * (sc / np.where(self.mh_u > 0, self.mh_u, 1.0)), 0.0)
            h_v = np.where(self.mh_v > 0, +e_u * (sc / np.where(self.mh_v > 0, self.mh_v, 1.0)), 0.0)
            Z, gam = z_mode(kind, omega_calc, wc, eps_r, z_line)
            zr = float(Z.real)
            v_test = float(np.dot(self.me_u, e_u**2)) + float(np.dot(self.me_v, e_v**2))
            i_test = float(np.dot(self.mh_u, h_u**2)) + float(np.dot(self.mh_v, h_v**2))
            hvp = h_v * self.mh_v / (MU0 * ndx)
            hup = h_u * self.mh_u / (MU0 * ndx)
            p_one = float(np.dot(e_u * hvp, area_u)) - float(np.dot(e_v * hup, area_v))
            g = (v_test / i_test) / zr
            kappa = math.sqrt(abs(p_one) * abs(zr))
            Hv = h_v * self.mh_v / (MU0 * ndx * pl["u_len"])
            Hu = h_u * self.mh_u / (MU0 * ndx * pl["v_len"])
            s_h = float(np.dot(e_u * Hv, area_u / pl["u_len"])) - float(np.dot(e_v * Hu, area_v / pl["v_len"]))
            hq = g * MU0 * ndx * (float(np.dot(h_v * Hv, pl["u_len"])) + float(np.dot(h_u * Hu, pl["v_len"])))
            c_state = math.sqrt(s_h / hq) / kappa
            self.state_scale[mi] = c_state
            self.record_scale[mi] = kappa * c_state
            self.source_scale[mi] = math.sqrt(abs(zr)) / kappa / c_state
            self.modes.append(dict(e_u=e_u, e_v=e_v, h_u=h_u * g, h_v=h_v * g, wc=wc, eps_r=eps_r,
                                   kind=kind, z_line=z_line))
        # chain parameters (all channels certified on a uniform feed)
        self.dtbc = []
        self.line = {}
        for mi, md in enumerate(self.modes):
            pair = np.concatenate([self.me_u * self.mh_v, self.me_v * self.mh_u])
            w = np.concatenate([self.me_u * md["e_u"] ** 2, self.me_v * md["e_v"] ** 2])
            act = w > 1e-12 * w.max()
            rp = dt / np.sqrt(pair[act])
            r = float(np.dot(w[act], rp) / w[act].sum())
            q = md["wc"] * dt
            num = np.concatenate([self.mh_v * md["h_v"], self.mh_u * md["h_u"]])
            den = np.concatenate([md["e_u"], md["e_v"]])
            cp = float(np.dot(w[act], np.abs(num[act] / den[act])) / w[act].sum())
            nv = float(np.dot(self.me_u, md["e_u"] ** 2) + np.dot(self.me_v, md["e_v"] ** 2))
            ni = float(np.dot(self.mh_u, md["h_u"] ** 2) + np.dot(self.mh_v, md["h_v"] ** 2))
            z0 = r * nv * cp / (dt * ni)
            if termination == "mur":
                self.dtbc.append(None)
            else:
                self.dtbc.append(DTBC(r, q) if kernel_init is None
                                 else DTBC(r, q, n_kernel_init=kernel_init))
                # no static z0 on a massless (q = 0) chain: its S-parameters
                # take the channel's own Z (operator.py:666-668)
                self.line[mi] = (r, q, z0 if q > 0.0 else None)
        self.mur_r = np.empty(M)
        self.tau = np.empty(M)
        for mi, md in enumerate(self.modes):
            _, gam = z_te(omega_calc, md["wc"], eps_r)
            beta = abs(gam.imag)
            vp = omega_calc / beta if beta > 0 else C0
            self.mur_r[mi] = (vp * dt - ndx) / (vp * dt + ndx)
            self.tau[mi] = ndx / vp
        self.src_buffer = collections.deque(maxlen=0)
        self.V_port_prev = np.zeros(M)
        self.absmax = 0.0  # |V| envelope since the last poll (operator.py:1530-1536)
        self.V_int_prev = np.zeros(M)
        self.exc = None
        self.wf = None
        self.V = [[] for _ in range(M)]
        self.I = [[] for _ in range(M)]

    @property
    def n_modes(self):
        return len(self.modes)

    def set_excitation(self, mode, waveform):
        s = float(self.source_scale[mode])
        self.exc = mode
        self.wf = (lambda t, _w=waveform, _s=s: _s * _w(t))
        buf_len = max(int(math.ceil((self.tau[mode] + self.dt) / self.dt)) + 3, 5)
        self.src_buffer = collections.deque(maxlen=buf_len)

    def _V(self, e, iu, iv, mu, mv):
        eu, ev = e[iu], e[iv]
        return np.array([float(np.dot(mu, md["e_u"] * eu)) + float(np.dot(mv, md["e_v"] * ev))
                         for md in self.modes])

    @property
    def complement_active(self):
        return self.comp is not None and any(term is None for term in self.dtbc)

    def update_e(self, e, t):
        pl = self.pl
        V_int_new = self._V(e, pl["e_u_int"], pl["e_v_int"], self.me_ui, self.me_vi)
        comp_on = self.complement_active
        if comp_on:
            cu = np.asarray(e[pl["e_u_int"]], float).copy()
            cv = np.asarray(e[pl["e_v_int"]], float).copy()
            for m, md in enumerate(self.modes):
                cu -= V_int_new[m] * md["e_u"]
                cv -= V_int_new[m] * md["e_v"]
        V = self.V_int_prev + self.mur_r * (V_int_new - self.V_port_prev)
        if self.exc is not None and self.dtbc[self.exc] is None:
            m = self.exc
            s_new = float(self.wf(t))
            self.src_buffer.append(s_new)
            tau, dt = self.tau[m], self.dt
            sin_now = V_int_new[m] - interp_delayed(self.src_buffer, tau, dt)
            sip = self.V_int_prev[m] - interp_delayed(self.src_buffer, tau + dt, dt)
            sfp = self.V_port_prev[m] - interp_delayed(self.src_buffer, dt, dt)
            V[m] = s_new + sip + self.mur_r[m] * (sin_now - sfp)
        for m, term in enumerate(self.dtbc):
            if term is None:
                continue
            src = float(self.wf(t - self.dt)) if m == self.exc else 0.0
            V[m] = term.advance(float(self.V_int_prev[m]), src)
        eu = np.zeros(pl["e_u"].size)
        ev = np.zeros(pl["e_v"].size)
        for m, md in enumerate(self.modes):
            eu += V[m] * md["e_u"]
            ev += V[m] * md["e_v"]
        if comp_on:
            r_u, r_v, live_u, live_v = self.comp
            pu = live_u * (self.comp_int_u + r_u * (cu - self.comp_port_u))
            pv = live_v * (self.comp_int_v + r_v * (cv - self.comp_port_v))
            eu += pu
            ev += pv
            self.comp_int_u, self.comp_int_v = cu, cv
            self.comp_port_u, self.comp_port_v = pu, pv
        e[pl["e_u"]] = eu
        e[pl["e_v"]] = ev
        self.V_port_prev[:] = V
        self.V_int_prev[:] = V_int_new
        env = float(np.max(np.abs(V) * self.record_scale))
        if env > self.absmax:
            self.absmax = env

    def poll_signal_absmax(self) -> float:
        """operator.py:1323-1334: the envelope since the last poll, reset."""
        m, self.absmax = self.absmax, 0.0
        return m

    def record(self, e, h):
        pl = self.pl
        Vp = self._V(e, pl["e_u"], pl["e_v"], self.me_u, self.me_v) * self.record_scale
        hu, hv = h[pl["h_u"]], h[pl["h_v"]]
        Ip = np.array([float(np.dot(self.mh_u, md["h_u"] * hu)) + float(np.dot(self.mh_v, md["h_v"] * hv))
                       for md in self.modes]) * self.record_scale
        for m in range(self.n_modes):
            self.V[m].append(float(Vp[m]))
            self.I[m].append(float(Ip[m]))


def complement_arrays(pl, m_eps, m_eps_vac, pec_e, dt):
    """(r_u, r_v, live_u, live_v) of the complement absorber on the plane's
    edges: ``m_eps`` / ``m_eps_vac`` flattened at the port face, ``pec_e``
    the mesh's own PEC edge mask (factory.py:687-735)."""
    out_r, out_dead = [], []
    for key in ("e_u", "e_v"):
        idx = pl[key]
        me = np.asarray(m_eps[idx], float)
        dead = me <= 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            c = C0 / np.sqrt(me / np.asarray(m_eps_vac[idx], float))
            r = (c * dt - pl["normal_dx"]) / (c * dt + pl["normal_dx"])
        out_r.append(np.where(dead, 0.0, r))
        out_dead.append(dead)
    live = [(~(np.asarray(pec_e, bool)[pl[k]] | d)).astype(float) for k, d in zip(("e_u", "e_v"), out_dead)]
    return out_r[0], out_r[1], live[0], live[1]


def interp_delayed(buffer: collections.deque, delay: float, dt: float) -> float:
    n = len(buffer)
    if n == 0 or dt <= 0:
        return 0.0
    frac = delay / dt
    i0 = int(math.floor(frac))
    a = frac - i0
    if n - 1 - i0 < 0:
        return 0.0
    v0 = buffer[n - 1 - i0]
    v1 = buffer[n - 2 - i0] if n - 2 - i0 >= 0 else 0.0
    return (1.0 - a) * v0 + a * v1
