"""Pins for the oracle's port model, and product-vs-oracle port setup (CPU).

* plane index maps: product and oracle agree bit for bit on all six faces
  of a non-uniform grid (integer work);
* the closed-form discrete TE modes are eigenpairs of the 2-D curl-curl
  pencil assembled here from the oracle's own 3-D curl operators
  (K e = w_c^2 M e to 1e-10), M_eps-orthonormal and M-orthogonal to node
  gradients (divergence-free);
* calibration identity: after the rescale the unit travelling wave has
  V/I = Re Z_modal exactly;
* chain parameters on a uniform feed: r = c dt / dz, q = w_c dt;
* the product's eigen-solver modes, scales and chain parameters agree with
  the oracle's closed forms (1e-9 relative).
"""

import math

import numpy as np
import pytest

import cases
from oracle import mesh as om
from oracle import ports as op
from oracle import yee


@pytest.mark.parametrize("face", ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
def test_plane_maps_bit_exact(face):
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    rng = np.random.default_rng(3)
    axes = [np.cumsum(np.r_[0, rng.uniform(0.5e-3, 1e-3, n)]) for n in (5, 6, 7)]
    pl_o = op.plane(face, *axes)
    pl_p = pp.PortPlane.from_mesh(face, mm.GridLines(*axes))
    for ko, kp in [("e_u", "e_u_indices"), ("h_v", "h_v_indices"), ("e_v", "e_v_indices"),
                   ("h_u", "h_u_indices"), ("e_u_int", "e_u_indices_interior"),
                   ("e_v_int", "e_v_indices_interior")]:
        assert np.array_equal(pl_o[ko], getattr(pl_p, kp)), ko
    assert np.array_equal(pl_o["u_uv"], pl_p.u_edge_uv)
    assert np.array_equal(pl_o["v_uv"], pl_p.v_edge_uv)
    assert np.array_equal(pl_o["u_len"], pl_p.u_edge_lengths)
    assert pl_o["normal_dx"] == pl_p.normal_dx


def _c1_masses():
    x, y, z = cases.c1_axes((12, 6, 10), (22.86e-3, 10.16e-3, 10e-3))
    mid = np.zeros((12, 6, 10), np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm_ = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    for f in ("zmin", "zmax"):
        me = om.flatten_port(me, 12, 6, 10, f, "E")
        mm_ = om.flatten_port(mm_, 12, 6, 10, f, "H")
    return (x, y, z), me, mm_


@pytest.mark.parametrize("face", ["zmin", "zmax"])
@pytest.mark.parametrize("mn", [(1, 0), (2, 0), (0, 1), (1, 1)])
def test_closed_form_modes_are_pencil_eigenpairs(face, mn):
    (x, y, z), me, mm_ = _c1_masses()
    dims = (12, 6, 10)
    pl = op.plane(face, x, y, z)
    e_u, e_v, wc = op.rect_te_mode(pl, x, y, z, *mn, me)
    e = np.zeros(yee.n_e(*dims))
    e[pl["e_u"]] = e_u
    e[pl["e_v"]] = e_v
    # h = C e on the normal faces of the plane only
    h = np.zeros(yee.n_h(*dims))
    yee.update_h(e, h, np.zeros_like(h), -np.ones_like(h), *dims)
    Hx, Hy, Hz = yee.split(h, *dims, "H")
    keep = np.zeros_like(h, dtype=bool)
    kz = 0 if face == "zmin" else dims[2]
    kx = yee.split(keep, *dims, "H")[2]
    kx[:, :, kz] = True
    h = np.where(keep, h / np.where(mm_ > 0, mm_, 1.0), 0.0)
    Ke = np.zeros_like(e)
    yee.update_e(Ke, h, np.zeros_like(e), np.ones_like(e), *dims)
    # the pencil acts on the free edges (PEC wall edges are Dirichlet rows)
    walls = om.bbox_walls(np.zeros(e.size, bool), *dims, ["xmin", "xmax", "ymin", "ymax"])
    free = ~np.concatenate([walls[pl["e_u"]], walls[pl["e_v"]]])
    lhs = np.concatenate([Ke[pl["e_u"]], Ke[pl["e_v"]]])[free]
    rhs = wc**2 * np.concatenate([me[pl["e_u"]] * e_u, me[pl["e_v"]] * e_v])[free]
    assert np.max(np.abs(lhs - rhs)) <= 1e-10 * np.max(np.abs(rhs))
    assert np.all(np.concatenate([e_u, e_v])[~free] == 0.0)
    assert float(np.sum(me[pl["e_u"]] * e_u**2) + np.sum(me[pl["e_v"]] * e_v**2)) == pytest.approx(1.0, abs=1e-13)
    # M-orthogonal to the gradient of any node potential vanishing on the walls
    rng = np.random.default_rng(1)
    phi = np.zeros((13, 7))
    phi[1:-1, 1:-1] = rng.standard_normal((11, 5))
    gx = phi[1:, :] - phi[:-1, :]
    gy = phi[:, 1:] - phi[:, :-1]
    g = np.zeros(yee.n_e(*dims))
    Ex, Ey, _ = yee.split(g, *dims, "E")
    Ex[:, :, kz] = gx
    Ey[:, :, kz] = gy
    assert abs(float(np.sum(me * e * g))) < 1e-12 * float(np.sqrt(np.sum(me * g * g)))


def test_calibration_identity_and_chain_params():
    orc, ports, dt, (me, mm_) = cases.c1_oracle(np.float64, dims=(12, 6, 10), size=(22.86e-3, 10.16e-3, 10e-3))
    p = ports[0]
    omega = 2 * math.pi * cases.C1_FMAX
    ndx = p.pl["normal_dx"]
    for md in p.modes:
        Z, _ = op.z_te(omega, md["wc"])
        # the unit travelling wave: dual voltages h = (+/-) e MU0 ndx / (Z M_mu)
        sc = op.MU0 * ndx / Z.real
        h_u_wave = -md["e_v"] * sc / p.mh_u
        h_v_wave = md["e_u"] * sc / p.mh_v
        v = float(np.dot(p.me_u, md["e_u"] ** 2) + np.dot(p.me_v, md["e_v"] ** 2))
        i = float(np.dot(p.mh_u, md["h_u"] * h_u_wave) + np.dot(p.mh_v, md["h_v"] * h_v_wave))
        assert v / i == pytest.approx(Z.real, rel=1e-12)
    dz = 10e-3 / 10
    for m, (r, q, z0) in p.line.items():
        assert r == pytest.approx(op.C0 * dt / dz, rel=1e-12)
        assert q == pytest.approx(p.modes[m]["wc"] * dt, rel=1e-15)
    assert np.all(p.record_scale > 0) and np.all(p.source_scale > 0)


def test_product_port_setup_matches_oracle():
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    dims, size = (12, 6, 10), (22.86e-3, 10.16e-3, 10e-3)
    orc, oports, dt, (me_o, mm_o) = cases.c1_oracle(np.float64, dims=dims, size=size)
    g = mm.GridLines(*cases.c1_axes(dims, size))
    mesh = mm.Mesh.from_grid(g, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    assert np.array_equal(me, om.mass_e(*cases.c1_axes(dims, size), np.zeros(dims, np.int32), [[1.0]] * 3, om.EPS0))
    for name, face, o in (("port1", "zmin", oports[0]), ("port2", "zmax", oports[1])):
        p = pp.build_port(name, face, mesh, me, mu, dt=dt, f_calc=cases.C1_FMAX, n_modes=3)
        assert p.termination_kinds == ["dtbc"] * 3
        for pm, om_ in zip(p.modes, o.modes):
            assert pm.omega_c == pytest.approx(om_["wc"], rel=1e-10)
            se = max(np.max(np.abs(om_["e_u"])), np.max(np.abs(om_["e_v"])))
            sh = max(np.max(np.abs(om_["h_u"])), np.max(np.abs(om_["h_v"])))
            for a, b, sc in ((pm.e_u, om_["e_u"], se), (pm.e_v, om_["e_v"], se),
                             (pm.h_u, om_["h_u"], sh), (pm.h_v, om_["h_v"], sh)):
                assert np.allclose(a, b, rtol=0, atol=1e-9 * sc)
        assert np.allclose(p.record_scale, o.record_scale, rtol=1e-9)
        assert np.allclose(p.source_scale, o.source_scale, rtol=1e-9)
        for m, (r, q, z0) in o.line.items():
            assert p.dtbc_r[m] == pytest.approx(r, rel=1e-12)
            assert p.dtbc_q[m] == pytest.approx(q, rel=1e-9)
            assert p.dtbc_z0[m] == pytest.approx(z0, rel=1e-9)


def _node_gradient_on_plane(phi, dims, kz):
    """[Ex | Ey] of the plane-node gradient of phi at node layer kz, as a
    flat E vector (the 3-D gradient of a potential living on one plane)."""
    g = np.zeros(yee.n_e(*dims))
    Ex, Ey, _ = yee.split(g, *dims, "E")
    Ex[:, :, kz] = phi[1:, :] - phi[:-1, :]
    Ey[:, :, kz] = phi[:, 1:] - phi[:, :-1]
    return g


@pytest.mark.parametrize("face", ["zmin", "zmax"])
@pytest.mark.parametrize("mn", [(1, 1), (2, 1), (1, 2), (3, 2)])
def test_closed_form_tm_modes_are_normal_e_pencil_eigenpairs(face, mn):
    """K_z phi = w_c^2 M_z phi with K_z applied through the oracle's own 3-D
    curls (normal edges of the first cell -> their transversal H faces ->
    back), the profile is the plane gradient of phi, curl-free on the plane,
    M_eps-normalised and M_eps-orthogonal to the TE modes."""
    (x, y, z), me, mm_ = _c1_masses()
    dims = nx, ny, nz = (12, 6, 10)
    pl = op.plane(face, x, y, z)
    m, n = mn
    e_u, e_v, wc = op.rect_tm_mode(pl, x, y, z, m, n, me)
    phi = np.outer(np.sin(m * math.pi * np.arange(nx + 1) / nx), np.sin(n * math.pi * np.arange(ny + 1) / ny))
    k_e = 0 if face == "zmin" else nz - 1  # normal-edge layer next to the plane
    e = np.zeros(yee.n_e(*dims))
    yee.split(e, *dims, "E")[2][:, :, k_e] = phi
    h = np.zeros(yee.n_h(*dims))
    yee.update_h(e, h, np.zeros_like(h), -np.ones_like(h), *dims)
    keep = np.zeros_like(h, dtype=bool)
    kx, ky, _ = yee.split(keep, *dims, "H")
    kx[:, :, k_e] = True
    ky[:, :, k_e] = True
    h = np.where(keep, h / np.where(mm_ > 0, mm_, 1.0), 0.0)
    Ke = np.zeros_like(e)
    yee.update_e(Ke, h, np.zeros_like(e), np.ones_like(e), *dims)
    lhs = yee.split(Ke, *dims, "E")[2][1:-1, 1:-1, k_e]
    rhs = wc**2 * (yee.split(me, *dims, "E")[2][:, :, k_e] * phi)[1:-1, 1:-1]
    assert np.max(np.abs(lhs - rhs)) <= 1e-10 * np.max(np.abs(rhs))
    # the transverse profile is the plane gradient of phi, normalised and signed
    kz = 0 if face == "zmin" else nz
    g = _node_gradient_on_plane(phi, dims, kz)
    full = np.concatenate([g[pl["e_u"]], g[pl["e_v"]]])
    mfull = np.concatenate([me[pl["e_u"]], me[pl["e_v"]]])
    full = op.resolve_sign(pl, full / math.sqrt(float(np.sum(mfull * full * full))))
    sc = float(np.abs(full).max())
    np.testing.assert_allclose(np.concatenate([e_u, e_v]), full, rtol=0, atol=1e-13 * sc)
    # curl-free: no normal H on the plane faces
    e_t = np.zeros_like(e)
    e_t[pl["e_u"]], e_t[pl["e_v"]] = e_u, e_v
    hn = np.zeros(yee.n_h(*dims))
    yee.update_h(e_t, hn, np.zeros_like(hn), -np.ones_like(hn), *dims)
    assert np.max(np.abs(yee.split(hn, *dims, "H")[2][:, :, kz])) <= 1e-13 * sc
    # M_eps-orthogonal to every TE mode of the plane
    for te in [(1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]:
        t_u, t_v, _ = op.rect_te_mode(pl, x, y, z, *te, me)
        ov = float(np.sum(me[pl["e_u"]] * t_u * e_u) + np.sum(me[pl["e_v"]] * t_v * e_v))
        assert abs(ov) <= 1e-12, te


def test_product_tm_modes_match_closed_form():
    """The product's TM eigen-solver on the uniform guide: the three lowest
    TM modes equal the closed-form TM_11, TM_21, TM_31 (cut-off 1e-10,
    profiles 1e-9 of their scale); the merged port keeps TE below TM."""
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp

    dims, size = (12, 6, 10), (22.86e-3, 10.16e-3, 10e-3)
    x, y, z = cases.c1_axes(dims, size)
    mesh = mm.Mesh.from_grid(mm.GridLines(x, y, z), pec_faces=("xmin", "xmax", "ymin", "ymax"))
    (_, _, _), me_o, _ = _c1_masses()
    for face in ("zmin", "zmax"):
        me = mm.flatten_port_mass(mm.build_m_eps(mesh), dims, face)
        mu = mm.flatten_port_mu(mm.build_m_mu(mesh), dims, face)
        pec = mm.pec_flat(mm.flatten_port_pec(mesh.pec_mask_edges, dims, face), dims)
        plane = pp.PortPlane.from_mesh(face, mesh.grid)
        got = pp.solve_tm_modes(plane, mesh.grid, me, mu, pec, 3, cases.C1_FMAX)
        pl = op.plane(face, x, y, z)
        for md, mn in zip(got, [(1, 1), (2, 1), (3, 1)]):
            e_u, e_v, wc = op.rect_tm_mode(pl, x, y, z, *mn, me_o)
            assert md.mode_type == "TM"
            assert md.omega_c == pytest.approx(wc, rel=1e-10)
            sc = max(np.abs(e_u).max(), np.abs(e_v).max())
            np.testing.assert_allclose(md.e_u, e_u, rtol=0, atol=1e-9 * sc)
            np.testing.assert_allclose(md.e_v, e_v, rtol=0, atol=1e-9 * sc)
        merged = pp.solve_modes(plane, mesh.grid, me, mu, pec, 6, cases.C1_FMAX)
        wcs = [md.omega_c for md in merged]
        assert wcs == sorted(wcs) and "TM" in [md.mode_type for md in merged]
