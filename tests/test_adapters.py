"""The reference's plug-in objects drive this backend (CPU side).

Stand-ins carry the reference's attribute names and, for the modal port,
the values a real ``PortOperatorModal`` held (tests/golden/ports_*.npz,
dumped from the reference's build_modal_port on a uniform and a graded
WR-90 cross-section and on an L-shaped section with a TM channel).  The adapters must rebuild exactly what this
backend's own builders make from the same mesh, and the solver fields must
carry over; ``resolve_backend`` serves "hip" only and never falls back.
"""

import math
from pathlib import Path
from types import SimpleNamespace as NS

import numpy as np
import pytest

from magnelio_amd import adapters, backend
from magnelio_amd import mesh as mm
from magnelio_amd import ports as pp
from magnelio_amd.boundaries import CPMLBoundary, PECBoundary
from magnelio_amd.monitors import MonitorFarField, MonitorFieldFrequency

GOLDEN = Path(__file__).resolve().parent / "golden"


def _operator_standin(f, label):
    """A PortOperatorModal-shaped object from the dumped attributes."""
    o = f"{label}_op_"
    g = {k[len(o):]: f[k] for k in f if k.startswith(o)}
    plane = NS(face=NS(value=str(g["face"])), coordinate=float(g["plane_coordinate"]),
               normal_dx=float(g["plane_normal_dx"]),
               **{k: g["plane_" + k] for k in ("e_u_indices", "h_v_indices", "u_edge_uv", "u_edge_lengths",
                                              "e_v_indices", "h_u_indices", "v_edge_uv", "v_edge_lengths",
                                              "e_u_indices_interior", "e_v_indices_interior",
                                              "u_node_window", "v_node_window")})
    modes = []
    for m in range(int(f.get("n_modes", 3))):
        mode = NS(name=str(g[f"mode{m}_name"]), mode_type=NS(value=str(g[f"mode{m}_type"])),
                  omega_c=float(g[f"mode{m}_omega_c"]), epsilon_r=float(g[f"mode{m}_epsilon_r"]))
        modes.append(NS(mode=mode, e_u_profile=f[f"{label}_e_u{m}"], e_v_profile=f[f"{label}_e_v{m}"],
                        h_u_profile=f[f"{label}_h_u{m}"], h_v_profile=f[f"{label}_h_v{m}"]))
    return NS(name=label, plane=plane, discrete_modes=modes, _dt=float(g["_dt"]),
              record_scale=f[f"{label}_record_scale"], _source_scale=f[f"{label}_source_scale"],
              state_scale=f[f"{label}_state_scale"], _comp_r_u=None,
              _chain_slab_defect=float(g["_chain_slab_defect"]),
              **{k: g[k] for k in ("_me_u_port", "_me_v_port", "_me_u_int", "_me_v_int", "_mh_u", "_mh_v",
                                   "_mur_r", "_tau_m", "_dtbc_r", "_dtbc_q", "_dtbc_z0",
                                   "_dtbc_pair_spread")})


@pytest.mark.parametrize("name", ["wr90", "graded", "corner"])
def test_port_from_operator_equals_own_build(name):
    f = dict(np.load(GOLDEN / f"ports_{name}.npz"))
    g = mm.GridLines(f["x"], f["y"], f["z"])
    # corner: an L-shaped section whose fourth channel is a TM mode
    corner = (0.0, 0.0, 0.0, 0.25 * 22.86e-3, 0.25 * 10.16e-3, 20e-3)
    regions = [(mm.Material.pec(), corner)] if bool(f.get("corner", False)) else []
    mesh = mm.Mesh.from_grid(g, regions=regions, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    dt = mm.courant_dt(g, "normal")
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    for label, face in (("port1", "zmin"), ("port2", "zmax")):
        own = pp.build_port(label, face, mesh, me, mu, dt=dt, f_calc=25e9, n_modes=int(f.get("n_modes", 3)))
        got = adapters.port_from_operator(_operator_standin(f, label), f_calc=25e9)
        assert [m.mode_type for m in got.modes] == [m.mode_type for m in own.modes]
        assert got.face == own.face == face
        for k in ("e_u_indices", "e_v_indices", "h_u_indices", "h_v_indices", "e_u_indices_interior",
                  "e_v_indices_interior"):
            assert np.array_equal(getattr(got.plane, k), getattr(own.plane, k)), k
        assert got.plane.normal_dx == own.plane.normal_dx
        for k in ("me_u_port", "me_v_port", "me_u_int", "me_v_int", "mh_u", "mh_v", "record_scale",
                  "source_scale", "state_scale", "mur_r", "tau"):
            np.testing.assert_allclose(getattr(got, k), getattr(own, k), rtol=1e-12, atol=0, err_msg=k)
        for k in ("dtbc_r", "dtbc_q", "dtbc_z0"):
            np.testing.assert_allclose(getattr(got, k), getattr(own, k), rtol=1e-12, err_msg=k)
        assert got.termination_kinds == own.termination_kinds
        for a, b in zip(got.modes, own.modes):
            assert a.omega_c == pytest.approx(b.omega_c, rel=1e-12)
            sc = max(np.abs(b.e_u).max(), np.abs(b.e_v).max())
            np.testing.assert_allclose(a.e_u, b.e_u, atol=1e-12 * sc)
            hs = max(np.abs(b.h_u).max(), np.abs(b.h_v).max())
            np.testing.assert_allclose(a.h_u, b.h_u, atol=1e-11 * hs)
            np.testing.assert_allclose(a.h_v, b.h_v, atol=1e-11 * hs)
        # the packed DTBC kernels the device reads are the same (their first
        # coefficient is FFT round-off, ~1e-17: gate on the kernel's scale)
        kg, ko = got.packed_dtbc_kernels(600), own.packed_dtbc_kernels(600)
        assert kg.shape == ko.shape and np.abs(kg - ko).max() <= 1e-12 * np.abs(ko).max()


def test_face_keys():
    assert adapters.face_key("z_min") == "zmin"
    assert adapters.face_key(NS(value="x_max")) == "xmax"
    assert adapters.face_key("ymin") == "ymin"
    with pytest.raises(ValueError):
        adapters.face_key("top")


def _ref_mesh():
    x, y, z = (np.linspace(0, s, n + 1) for s, n in ((8e-3, 6), (5e-3, 4), (9e-3, 7)))
    ref = mm.Mesh.from_grid(mm.GridLines(x, y, z),
                            regions=[(mm.Material.dielectric(3.0, sigma=0.2), (0, 0, 3e-3, 8e-3, 5e-3, 6e-3)),
                                     (mm.Material.pec(), (3e-3, 0, 4e-3, 5e-3, 5e-3, 5e-3))],
                            pec_faces=("xmin", "xmax"))
    disp = NS(poles=((complex(-1e9, 2e10), complex(0, -3e10)),))
    lib = {k: NS(name=m.name, epsilon=m.epsilon, mu=m.mu, sigma=m.sigma, sigma_m=m.sigma_m, is_pec=m.is_pec,
                 dispersion=disp if k == 1 else None, dispersion_mu=None)
           for k, m in ref.material_library.items()}
    stand = NS(grid=NS(x=ref.grid.x, y=ref.grid.y, z=ref.grid.z), material_id=ref.material_id,
               material_library=lib, pec_mask_edges=ref.pec_mask_edges.copy(),
               boundary_conditions={"xmin": "PEC", "xmax": NS(kind="PEC"), "ymin": "CPML", "ymax": "PMC",
                                    "zmin": "CPML", "zmax": "CPML"})
    return ref, stand


def test_mesh_from_reference():
    own, stand = _ref_mesh()
    got = adapters.mesh_from_reference(stand)
    assert np.array_equal(got.material_id, own.material_id)
    assert np.array_equal(got.pec_mask_edges, own.pec_mask_edges)
    assert got.pec_faces == ("xmin", "xmax") and not got.pec_rule
    assert np.array_equal(mm.build_m_eps(got), mm.build_m_eps(own))
    assert np.array_equal(mm.build_m_sigma(got), mm.build_m_sigma(own))
    assert got.material_library[1].poles == ((complex(-1e9, 2e10), complex(0, -3e10)),)
    assert got.material_library[1].poles_mu == ()
    stand.material_library[1].dispersion_mu = NS(poles=((-2e9, 5e9),))
    got = adapters.mesh_from_reference(stand)
    assert got.material_library[1].poles_mu == ((complex(-2e9), complex(5e9)),)


def test_boundaries_sources_monitors():
    g = mm.GridLines(np.linspace(0, 1e-2, 11), np.linspace(0, 1e-2, 11), np.linspace(0, 1e-2, 11))
    ref_cpml = NS(face=NS(value="z_max"), thickness_cells=5, m=4, R_target=1e-6, kappa_max=3.0, alpha_max=0.05)
    c = adapters.boundary_from_reference("z_max", ref_cpml, g)
    assert isinstance(c, CPMLBoundary) and c.face == "zmax"
    assert (c.thickness_cells, c.m, c.R_target, c.kappa_max, c.alpha_max) == (5, 4, 1e-6, 3.0, 0.05)
    PECBoundaryRef = type("PECBoundary", (), {})
    assert isinstance(adapters.boundary_from_reference("x_min", PECBoundaryRef(), g), PECBoundary)
    src = adapters.plane_wave_from_reference(NS(direction=(0, 0, -1.0), polarization=(0, 1.0, 0),
                                                corners=((1e-3,) * 3, (9e-3,) * 3), amplitude=2.0,
                                                waveform="gaussian", f_center=None, f_max=3e10))
    assert src.direction == (0.0, 0.0, -1.0) and src.amplitude == 2.0 and src.f_max == 3e10
    region = adapters.monitor_from_reference(NS(freqs=[1e9, 2e9], corners=((None, None, 5e-3),) * 2,
                                                fields=["E"], interval=None, name="cut"), block=128)
    assert isinstance(region, MonitorFieldFrequency) and region.name == "cut" and region.block == 128
    assert region.components == ["Ex", "Ey", "Ez"]
    far = adapters.monitor_from_reference(NS(freqs=[1e9], margin_cells=2, name="far"))
    assert isinstance(far, MonitorFarField) and far.margin_cells == 2
    with pytest.raises(NotImplementedError, match="interval"):
        adapters.monitor_from_reference(NS(freqs=[1e9], corners=None, fields=["E"], interval=1e-11, name="d"))


def test_solver_from_reference_fields():
    own, stand = _ref_mesh()
    wf_calls = []

    def wf(t):
        wf_calls.append(t)
        return 3.0 * math.sin(t)

    sol = adapters.solver_from_reference(
        mesh=stand, boundary_conditions={"zmin": NS(face=NS(value="z_min"), thickness_cells=2, m=3,
                                                    R_target=1e-8, kappa_max=7.0, alpha_max=0.02)},
        total_time_steps=None, max_time_steps=500, dt=1e-13, energy_stop_db=60.0, port_signal_min_steps=7,
        energy_check_interval=20, precision="double", excitations=[None, None])
    assert sol.nsets == 2 and sol.precision == "double" and sol.horizon == 500
    assert sol.energy_stop_db == 60.0 and sol.port_signal_min_steps == 7 and sol.energy_check_interval == 20
    assert isinstance(sol.boundary_conditions["zmin"], CPMLBoundary)
    spec = NS(surfaces=(NS(tag="xmin", comp=np.array([2], np.uint8), flat_idx=np.array([0]),
                           weight=np.array([1e-6]), inv_l_dual=np.array([1e3])),),
              fits={"xmin": NS(c0=0.01, branches=((1e10, 0.02),))})
    sol = adapters.solver_from_reference(mesh=stand, total_time_steps=10, dt=1e-13, sibc=spec)
    assert sol.sibc is spec


def test_resolve_backend_never_falls_back(monkeypatch):
    with pytest.raises(ValueError, match="Unknown backend"):
        backend.resolve_backend("numpy")
    monkeypatch.setenv("MAGNELIO_BACKEND", "cupy")
    with pytest.raises(ValueError, match="not served"):
        backend.resolve_backend("auto")
    monkeypatch.delenv("MAGNELIO_BACKEND")
    import torch

    if not torch.cuda.is_available():  # this CPU container: a loud error, no CPU path
        backend._PROBED.clear()
        with pytest.raises(backend.HipDeviceUnavailable):
            backend.resolve_backend("hip")
        with pytest.raises(backend.HipDeviceUnavailable):
            backend.resolve_backend("auto")
        backend._PROBED.clear()


def test_recorder_signals_shape_of_finalize():
    """recorder_signals gives PortSignalRecorder.finalize's mapping: one
    (V, I) pair per (port, mode) over t = arange(N) dt, labelled
    "{port}_mode{m}_V" / "_I" (ports/recorder.py:319-340)."""
    from dataclasses import dataclass

    @dataclass(frozen=True)
    class Sig:
        t: np.ndarray
        values: np.ndarray
        dt: float
        label: str = ""

    V = np.arange(5.0)
    rec = {("port1", 0): (V, -V), ("port2", 1): (2 * V, V + 1)}
    sol = NS(dt=1.5e-12, recorded=lambda s: rec if s == 0 else {})
    assert adapters.recorder_signals(sol, 0) is rec
    out = adapters.recorder_signals(sol, 0, Sig)
    assert set(out) == set(rec)
    v, i = out[("port2", 1)]
    assert v.label == "port2_mode1_V" and i.label == "port2_mode1_I"
    assert np.array_equal(v.values, 2 * V) and np.array_equal(i.values, V + 1)
    assert np.allclose(v.t, np.arange(5) * 1.5e-12) and v.dt == 1.5e-12
