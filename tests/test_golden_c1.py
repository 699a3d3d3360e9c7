"""Whole-run golden vectors of the reference's own solver (CPU side).

``tests/golden/c1_{guide,loaded,mur,tm}.npz`` hold 300 steps of the
reference's NumPy tier on reduced C1 cases (tests/golden/make_c1_fixtures.py):
the uniform WR-90 guide, a graded cross-section with a lossy dielectric
block and a PEC post, a dielectric block that puts port 1 on Mur, and an
L-shaped section whose 4-mode ports carry a TM channel (the one excited).  They pin the oracle's whole composition — graded dual
masses, material averaging, the PEC edge OR rule, port-plane flattening,
the curl order, the modal DTBC port step and calibration, the recorder and
the S-parameter post — and the product's host coefficient build.

* product alpha/beta (host build) == the reference's, bit for bit;
* oracle alpha/beta (oracle.mesh) == the reference's, bit for bit;
* oracle float64 fields after 1, 10, 300 steps (1e-12; measured <= 2.3e-13,
  the port dot products associate differently), every recorded V/I and the
  S-parameters on 41 frequencies (1e-12; measured 2e-15) against the
  reference's;
* oracle float32 (the reference's order) equal to the reference's float32
  run of the guide bit for bit, every step snapshot.
"""

import numpy as np
import pytest

import cases
from magnelio_amd import mesh as pm
from oracle import sparams as osp

CASES = ("c1_guide", "c1_loaded", "c1_mur", "c1_tm", "c1_coax")


def _product_coefficients(name):
    mesh = cases.golden_c1_mesh(name)
    dims = mesh.dims
    f = cases.golden_c1(name)
    dt = pm.courant_dt(mesh.grid, "normal")
    assert dt == float(f["dt"])
    me, mu = pm.build_m_eps(mesh), pm.build_m_mu(mesh)
    pec = mesh.pec_mask_edges
    for face in ("zmin", "zmax"):
        me = pm.flatten_port_mass(me, dims, face)
        mu = pm.flatten_port_mu(mu, dims, face)
        pec = pm.flatten_port_pec(pec, dims, face)
    return pm.update_coefficients(me, mu, pm.build_m_sigma(mesh), pm.build_m_sigma_m(mesh), dt,
                                  pm.pec_flat(pec, dims))


@pytest.mark.parametrize("name", CASES)
def test_product_coefficients_equal_reference(name):
    f = cases.golden_c1(name)
    co = _product_coefficients(name)
    for k in ("alpha_e", "beta_e", "alpha_h", "beta_h"):
        assert np.array_equal(getattr(co, k), f["double_" + k]), k


@pytest.mark.parametrize("name", CASES)
def test_oracle_coefficients_equal_reference(name):
    f = cases.golden_c1(name)
    orc, _, _, _ = cases.golden_c1_oracle(name)
    for k, v in (("alpha_e", orc.ae), ("beta_e", orc.be), ("alpha_h", orc.ah), ("beta_h", orc.bh)):
        assert np.array_equal(v, f["double_" + k]), k


def _sparams(ports, dt, f_axis, excited=("port1", 0)):
    sig = {(p.name, m): (np.array(p.V[m]), np.array(p.I[m])) for p in ports for m in range(p.n_modes)}
    lines = {(p.name, m): p.line[m] for p in ports for m in range(p.n_modes) if m in p.line}
    modes = {(p.name, m): (md["wc"], md["eps_r"], md["kind"], md["z_line"]) for p in ports
             for m, md in enumerate(p.modes)}
    ndx = {p.name: p.pl["normal_dx"] for p in ports}
    return osp.s_params(sig, lines, excited, f_axis, dt, modes=modes, normal_dx=ndx)


@pytest.mark.parametrize("name", CASES)
def test_oracle_run_matches_reference(name):
    f = cases.golden_c1(name)
    orc, ports, dt, _ = cases.golden_c1_oracle(name)
    # relative to the snapshot's norm, or to a tenth of the run's largest
    # one when the field has decayed below that (the coax pulse has left its
    # line by step 300: field 1e-12 of its peak)
    peak = {c: max(np.linalg.norm(f[f"double_{c}{n}"]) for n in f["snaps"]) for c in "eh"}
    for n in f["snaps"]:
        orc.run(int(n) - orc.n)
        for comp, got in (("e", orc.e), ("h", orc.h)):
            want = f[f"double_{comp}{n}"]
            err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 0.1 * peak[comp])
            assert err <= 1e-12, (n, comp, err)
    # channels that carry no mode stay at round-off: gate every channel
    # against the largest V (I) of the run
    vmax = max(np.abs(f[k]).max() for k in f if k.startswith("double_V_"))
    imax = max(np.abs(f[k]).max() for k in f if k.startswith("double_I_"))
    for p in ports:
        for m in range(p.n_modes):
            np.testing.assert_allclose(p.V[m], f[f"double_V_{p.name}_{m}"], rtol=0, atol=1e-10 * vmax)
            np.testing.assert_allclose(p.I[m], f[f"double_I_{p.name}_{m}"], rtol=0, atol=1e-10 * imax)
    S = _sparams(ports, dt, f["f_axis"], ("port1", cases.golden_excite(name)))
    for (lbl, m), s in S.items():
        want = f[f"double_S_{lbl}_{m}"]
        ok = np.isfinite(want)
        assert np.array_equal(ok, np.isfinite(s))
        assert np.max(np.abs(s[ok] - want[ok])) <= 1e-12, (lbl, m)


def test_complement_absorber_arrays_equal_reference():
    """The oracle's per-edge Mur coefficients and live masks of every port
    (factory.py:687-735) equal the reference's; c1_mur's port 1 is the one
    on Mur, with its absorber live."""
    f = cases.golden_c1("c1_mur")
    assert f["port1_mur"].all() and not f["port2_mur"].any()
    _, ports, _, _ = cases.golden_c1_oracle("c1_mur")
    for p in ports:
        for k, got in zip(("r_u", "r_v", "live_u", "live_v"), p.comp):
            np.testing.assert_allclose(got, f[f"{p.name}_comp_{k}"], rtol=1e-14, atol=0)
    assert ports[0].complement_active and not ports[1].complement_active


@pytest.mark.parametrize("name", ["c1_guide", "c1_mur"])
def test_product_ports_carry_reference_complement(name):
    """The product's port build: the reference's terminations (c1_mur's
    port 1 on Mur, everything else on the DTBC) and complement arrays."""
    f = cases.golden_c1(name)
    mesh = cases.golden_c1_mesh(name)
    dt = pm.courant_dt(mesh.grid, "normal")
    from magnelio_amd import ports as pp  # noqa: PLC0415

    me, mu = pm.build_m_eps(mesh), pm.build_m_mu(mesh)
    for label, face in (("port1", "zmin"), ("port2", "zmax")):
        P = pp.build_port(label, face, mesh, me, mu, dt=dt, f_calc=cases.C1_FMAX, n_modes=3)
        if f"{label}_mur" in f:
            assert [k == "mur" for k in P.termination_kinds] == list(f[f"{label}_mur"])
            np.testing.assert_allclose(P.mur_r, f[f"{label}_mur_r"], rtol=1e-12)
            n_u = P.plane.e_u_indices.size
            np.testing.assert_allclose(P.comp_r[:n_u], f[f"{label}_comp_r_u"], rtol=1e-14, atol=0)
            np.testing.assert_allclose(P.comp_r[n_u:], f[f"{label}_comp_r_v"], rtol=1e-14, atol=0)
            assert np.array_equal(P.comp_live, np.concatenate([f[f"{label}_comp_live_u"],
                                                               f[f"{label}_comp_live_v"]]))
        assert P.complement_active == (name == "c1_mur" and label == "port1")


def test_oracle_checkpoint_complement_state_equals_reference():
    f = cases.golden_c1("c1_mur")
    orc, ports, _, _ = cases.golden_c1_oracle("c1_mur")
    orc.run(150)
    p = ports[0]
    for k, got in (("int_u", p.comp_int_u), ("int_v", p.comp_int_v), ("port_u", p.comp_port_u),
                   ("port_v", p.comp_port_v)):
        want = f[f"double_state/ports/port1/complement/{k}"]
        assert np.max(np.abs(got - want)) <= 1e-12 * np.abs(want).max(), k


@pytest.mark.parametrize("name", ["c1_guide", "c1_mur", "c1_tm", "c1_coax"])
def test_oracle_float32_matches_reference_float32(name):
    f = cases.golden_c1(name)
    orc, _, _, _ = cases.golden_c1_oracle(name, np.float32)
    for n in f["snaps"]:
        orc.run(int(n) - orc.n)
        for comp, got in (("e", orc.e), ("h", orc.h)):
            want = f[f"single_{comp}{n}"]
            if name == "c1_coax":
                # the float64 port projections of the TEM profile differ from
                # the reference's NumPy dots in the last bit; once the pulse
                # has left the line (step 300) such a flip shows in a few
                # float32 words of the 1e-9 residue
                assert np.abs(got.astype(np.float64) - want).max() <= 1e-6 * np.abs(want).max(), (n, comp)
                continue
            # same operation order in float32: bit for bit
            assert np.array_equal(got.astype(np.float64), want), (n, comp)


def test_product_coax_port_matches_reference():
    """The product's coaxial port (conductors from the mesh, Laplace TEM
    channel, line impedance, calibration, q = 0 chain) on c1_coax equals the
    reference's PortSpecCoax build."""
    f = cases.golden_c1("c1_coax")
    mesh = cases.golden_c1_mesh("c1_coax")
    dt = pm.courant_dt(mesh.grid, "normal")
    me, mu = pm.build_m_eps(mesh), pm.build_m_mu(mesh)
    for label, face in (("port1", "zmin"), ("port2", "zmax")):
        P = cases.golden_port("c1_coax", label, face, mesh, me, mu, dt, 1)
        (md,) = P.modes
        assert md.mode_type == "TEM" and md.omega_c == 0.0
        assert md.z_line == pytest.approx(float(f[f"{label}_z_line0"]), rel=1e-12)
        sc = max(np.abs(f[f"{label}_e_u0"]).max(), np.abs(f[f"{label}_e_v0"]).max())
        np.testing.assert_allclose(md.e_u, f[f"{label}_e_u0"], rtol=0, atol=1e-12 * sc)
        np.testing.assert_allclose(md.e_v, f[f"{label}_e_v0"], rtol=0, atol=1e-12 * sc)
        for k in ("record_scale", "source_scale", "state_scale"):
            np.testing.assert_allclose(getattr(P, k), f[f"{label}_{k}"], rtol=1e-12, err_msg=k)
        r, q, z0 = f[f"{label}_dtbc"][0]
        assert P.dtbc_r[0] == pytest.approx(r, rel=1e-12) and P.dtbc_q[0] == q == 0.0
        assert P.dtbc_z0[0] is None and np.isnan(z0)
        np.testing.assert_allclose(P.mur_r, f[f"{label}_mur_r"], rtol=1e-12)
