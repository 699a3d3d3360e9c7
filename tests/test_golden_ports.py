"""Modal-port calibration against the reference's own values (golden vectors).

``tests/golden/ports_{wr90,graded,corner}.npz`` were written by
``tests/golden/make_port_fixtures.py`` from the reference's
``build_modal_port`` (src/magnelio/ports/_modal/factory.py:1367) on a WR-90
guide with a uniform and with a seeded graded cross-section, 3 numerical
modes per port on z_min / z_max, and on an L-shaped section (a PEC block in
one corner) whose 4-mode ports carry the lowest TM mode as channel 3.  They pin what no closed form fixes: the
V/I calibration ``record_scale`` / ``source_scale`` / ``state_scale``
(operator.py:778-941), the DTBC chain (r, q, z0) (operator.py:671-776), the
cut-offs and the M_eps-normalised mode profiles (ARPACK start-vector noise
~1e-13 of the profile scale).

* the product's port build (``magnelio_amd.ports.build_port``) on all
  three: mode families, cut-offs, E and calibrated H profiles, scales and
  chain parameters;
* the oracle's closed-form TE modes and calibration (``oracle.ports``) on
  the uniform guide, where its closed forms apply.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd import ports as pp
from oracle import mesh as om
from oracle import ports as op

GOLDEN = Path(__file__).resolve().parent / "golden"
PORTS = (("port1", "zmin"), ("port2", "zmax"))


def _load(name):
    return dict(np.load(GOLDEN / f"ports_{name}.npz"))


def _fixture_module():
    import sys  # noqa: PLC0415

    if str(GOLDEN) not in sys.path:
        sys.path.insert(0, str(GOLDEN))
    import make_port_fixtures as mk  # noqa: PLC0415

    return mk


@pytest.mark.parametrize("name", ["wr90", "graded", "corner"])
def test_product_port_build_matches_reference(name):
    f = _load(name)
    g = mm.GridLines(f["x"], f["y"], f["z"])
    regions = [(mm.Material.pec(), _fixture_module().CORNER)] if bool(f.get("corner", False)) else []
    mesh = mm.Mesh.from_grid(g, regions=regions, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    n_modes = int(f.get("n_modes", 3))
    dt = mm.courant_dt(g, "normal")
    assert dt == float(f["dt"])
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    for label, face in PORTS:
        P = pp.build_port(label, face, mesh, me, mu, dt=dt, f_calc=float(f["f_calc"]), n_modes=n_modes)
        p = label + "_"
        assert P.termination_kinds == [str(k) for k in f[p + "kinds"]]
        if p + "types" in f:
            assert [md.mode_type for md in P.modes] == [str(t) for t in f[p + "types"]]
        for key in ("record_scale", "source_scale", "state_scale"):
            np.testing.assert_allclose(getattr(P, key), f[p + key], rtol=1e-12, atol=0, err_msg=key)
        np.testing.assert_allclose([md.omega_c for md in P.modes], f[p + "omega_c"], rtol=1e-12)
        line = np.array([[P.dtbc_r[m], P.dtbc_q[m], P.dtbc_z0[m]] for m in range(n_modes)])
        np.testing.assert_allclose(line, f[p + "dtbc"], rtol=1e-12, atol=0)
        assert np.array_equal(P.plane.e_u_indices, f[p + "e_u_indices"])
        assert np.array_equal(P.plane.e_v_indices, f[p + "e_v_indices"])
        for m in range(n_modes):
            for fam in ("e", "h"):
                scale = max(np.abs(f[f"{p}{fam}_u{m}"]).max(), np.abs(f[f"{p}{fam}_v{m}"]).max())
                for comp in (fam + "_u", fam + "_v"):
                    np.testing.assert_allclose(getattr(P.modes[m], comp), f[f"{p}{comp}{m}"], rtol=0,
                                               atol=1e-12 * scale, err_msg=f"{label} mode {m} {comp}")


# (m, n) of the three lowest TE cut-offs of WR-90: TE10, TE20, TE01
WR90_MODES = [(1, 0), (2, 0), (0, 1)]


def test_oracle_calibration_matches_reference_on_uniform_guide():
    f = _load("wr90")
    x, y, z = f["x"], f["y"], f["z"]
    dims = (x.size - 1, y.size - 1, z.size - 1)
    mid = np.zeros(dims, np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mu = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    for face in ("zmin", "zmax"):
        me = om.flatten_port(me, *dims, face, "E")
        mu = om.flatten_port(mu, *dims, face, "H")
    dt = float(f["dt"])
    for label, face in PORTS:
        pl = op.plane(face, x, y, z)
        modes = [op.rect_te_mode(pl, x, y, z, m, n, me) for m, n in WR90_MODES]
        P = op.Port(label, pl, modes, me, mu, dt, 2 * math.pi * float(f["f_calc"]))
        p = label + "_"
        assert np.array_equal(pl["e_u"], f[p + "e_u_indices"])
        assert np.array_equal(pl["e_v"], f[p + "e_v_indices"])
        np.testing.assert_allclose([md["wc"] for md in P.modes], f[p + "omega_c"], rtol=1e-10)
        for key in ("record_scale", "source_scale", "state_scale"):
            np.testing.assert_allclose(getattr(P, key), f[p + key], rtol=1e-10, err_msg=key)
        line = np.array([P.line[m] for m in range(3)])
        np.testing.assert_allclose(line, f[p + "dtbc"], rtol=1e-10)
        for m in range(3):
            scale = max(np.abs(f[f"{p}e_u{m}"]).max(), np.abs(f[f"{p}e_v{m}"]).max())
            np.testing.assert_allclose(P.modes[m]["e_u"], f[f"{p}e_u{m}"], atol=1e-10 * scale)
            np.testing.assert_allclose(P.modes[m]["e_v"], f[f"{p}e_v{m}"], atol=1e-10 * scale)
