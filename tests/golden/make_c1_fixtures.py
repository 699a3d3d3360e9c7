"""Write golden vectors of whole FIT-TD runs from the reference's own solver.

Run in the build container only (the reference never travels to the GPU
machine; only the .npz files this writes do):

    python tests/golden/make_c1_fixtures.py /root/reference/src [case ...]

The reference's NumPy tier (``FITTimeDomainSolver``, src/magnelio/solver/
fit_td.py:682-1246, with ``PortOperatorModal`` ports, ``PortSignalRecorder``
and ``compute_s_parameters``, post/modal_sparameters.py:335) runs two reduced
plain cases for 300 steps:

* ``c1_guide``:  the C1 WR-90 guide on a 12 x 8 x 20 uniform grid (1 mm in
  z), air, PEC x/y walls, 3-mode numerical ports on z_min / z_max;
* ``c1_loaded``: the same guide on a seeded graded cross-section (adjacent
  spacing ratio <= 1.3) with a lossy dielectric block (eps_r 4, sigma
  0.05 S/m) and a PEC post in its middle — graded dual masses, material
  averaging, the PEC edge OR rule and the port-plane flattening all enter;
* ``c1_mur``:    the guide with a dielectric block (eps_r 2.5) in z cells 2-4
  next to port 1: its feed slabs fail the DTBC certificate, its channels run
  on modal Mur and the complement absorber (a per-edge Mur on the
  port-unrepresented remainder of the plane) is live; port 2 keeps the DTBC;
* ``c1_tm``:     the guide with a PEC block filling a 3 x 2-cell corner of the
  cross-section along its whole length (an L-shaped section: TE_mn and
  TM_mn are no longer degenerate), 4-mode ports whose channel 3 is the
  lowest TM mode; that TM channel of port 1 is the one excited;
* ``c1_coax``:   a square coaxial line, 12 x 12 x 20 cells of 1 mm: PEC box
  walls around a 4 x 4-cell PEC inner conductor along the whole length,
  coaxial (TEM) ports on z_min / z_max (``PortSpecCoax``);

port 1 mode 0 (c1_tm: mode 3) excited with the modulated Gaussian from its
cut-off to 25 GHz (the TEM channel of c1_coax: the plain Gaussian to
25 GHz), dt = courant_dt(grid, "normal"), float64 (and float32 for the
guide).  Stored: the ports' numerical mode profiles and cut-offs, alpha/beta
(E, H) of the solver, e/h after steps 1, 10, 150 and 300 (reference flat
layout), the solver's state_dict() after step 150 (flattened "a/b" keys),
the recorded V/I of every channel and the S-parameters of the excited column
on 41 frequencies.
"""

from __future__ import annotations

import math
import sys
import warnings
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from _flat import flatten  # noqa: E402
A, B = 22.86e-3, 10.16e-3
F_MAX = 25e9
SNAPS = (1, 10, 150, 300)
CHECKPOINT = 150  # the reference's state_dict() is stored after this step


def graded_axis(n, length, seed):
    rng = np.random.default_rng(seed)
    w = np.exp(np.concatenate([[0.0], np.cumsum(rng.uniform(np.log(1 / 1.3), np.log(1.3), n - 1))]))
    w = w / w.sum() * length
    return np.concatenate([[0.0], np.cumsum(w)])


def cases():
    z = np.linspace(0.0, 20e-3, 21)
    return {
        "c1_guide": dict(axes=(np.linspace(0.0, A, 13), np.linspace(0.0, B, 9), z), loaded=False,
                         precisions=("double", "single")),
        "c1_loaded": dict(axes=(graded_axis(12, A, 21), graded_axis(8, B, 22), z), loaded=True,
                          precisions=("double",)),
        # a dielectric block from the third cell at z_min: port 1's feed
        # slabs differ, its DTBC slab certificate fails and its channels
        # terminate on Mur with the complement absorber live (operator.py:
        # 596-619, :1337-1346); port 2 keeps the DTBC
        "c1_mur": dict(axes=(np.linspace(0.0, A, 13), np.linspace(0.0, B, 9), z), loaded="mur",
                       precisions=("double", "single")),
        # an L-shaped section: the lowest TM mode is channel 3 of 4
        "c1_tm": dict(axes=(np.linspace(0.0, A, 13), np.linspace(0.0, B, 9), z), loaded="tm",
                      precisions=("double", "single"), n_modes=4, excite=3),
        # a square coax: TEM ports
        "c1_coax": dict(axes=(np.linspace(0.0, COAX_W, 13), np.linspace(0.0, COAX_W, 13), z),
                        loaded="coax", precisions=("double", "single"), n_modes=1, excite=0),
    }


# the loaded case: dielectric block and PEC post (x0, y0, z0, x1, y1, z1)
BLOCK = (0.25 * A, 0.0, 8e-3, 0.75 * A, 0.6 * B, 12e-3)
POST = (0.45 * A, 0.0, 9e-3, 0.55 * A, B, 11e-3)
BLOCK_EPS, BLOCK_SIGMA = 4.0, 0.05
# the Mur case: a block in z cells 2-4 next to port 1
MUR_BLOCK = (0.3 * A, 0.0, 2e-3, 0.7 * A, 0.5 * B, 5e-3)
MUR_EPS = 2.5
# the TM case: a PEC block over cells x 0-2, y 0-1, the whole length
TM_CORNER = (0.0, 0.0, 0.0, 0.25 * A, 0.25 * B, 20e-3)
# the coax case: a 12 mm square box, inner conductor over cells 4-7 in x and y
COAX_W = 12e-3
COAX_INNER = (4e-3, 4e-3, 0.0, 8e-3, 8e-3, 20e-3)


def main(ref_src: str, only=None) -> None:
    sys.path.insert(0, ref_src)
    from magnelio.materials.material import Material
    from magnelio.mesh.grid import GridLines
    from magnelio.mesh.mesher import Mesh
    from magnelio._operators.material_matrices import build_M_eps, build_M_mu
    from magnelio.ports._modal import BoxFace, PortSpecCoax, PortSpecNumerical, build_modal_port
    from magnelio.ports.recorder import PortSignalRecorder
    from magnelio.post.modal_sparameters import compute_s_parameters
    from magnelio.signals.signal_1d import Signal1D
    from magnelio.signals.waveforms import gaussian, modulated_gaussian
    from magnelio.solver.fit_td import FITTimeDomainSolver
    from magnelio.solver.stability import courant_dt

    for name, cfg in cases().items():
        if only and name not in only:
            continue
        x, y, z = cfg["axes"]
        grid = GridLines(x=x, y=y, z=z)
        regions = []
        n_modes, excite = cfg.get("n_modes", 3), cfg.get("excite", 0)
        if cfg["loaded"] == "coax":
            regions = [(Material.pec(), COAX_INNER)]
        elif cfg["loaded"] == "tm":
            regions = [(Material.pec(), TM_CORNER)]
        elif cfg["loaded"] == "mur":
            regions = [(Material(name="feed_block", epsilon=(MUR_EPS,) * 3), MUR_BLOCK)]
        elif cfg["loaded"]:
            regions = [(Material(name="block", epsilon=(BLOCK_EPS,) * 3, sigma=(BLOCK_SIGMA,) * 3), BLOCK),
                       (Material.pec(), POST)]
        mesh = Mesh.from_grid(grid, regions=regions).with_boundary_conditions(
            {"xmin": "PEC", "xmax": "PEC", "ymin": "PEC", "ymax": "PEC", "zmin": "PMC", "zmax": "PMC"})
        dt = courant_dt(grid, "normal")
        out = {"x": x, "y": y, "z": z, "dt": np.float64(dt), "snaps": np.array(SNAPS),
               "n_modes": np.int64(n_modes), "excite": np.int64(excite)}
        for prec in cfg["precisions"]:
            m_eps, m_mu = build_M_eps(mesh), build_M_mu(mesh)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                def spec(lbl, face):
                    if cfg["loaded"] == "coax":
                        return PortSpecCoax(name=lbl, plane=face, inner_radius=2e-3, outer_radius=6e-3,
                                            center=(0.5 * COAX_W, 0.5 * COAX_W), n_modes=n_modes)
                    return PortSpecNumerical(name=lbl, plane=face, n_modes=n_modes)

                ops = [build_modal_port(spec(lbl, face), mesh, m_eps, m_mu, dt=dt, f_calc=F_MAX)
                       for lbl, face in (("port1", BoxFace.Z_MIN), ("port2", BoxFace.Z_MAX))]
            f_c = ops[0].discrete_modes[excite].mode.omega_c / (2 * math.pi)
            if prec == "double":  # the ports' numerical modes (the oracle takes them as given)
                for op in ops:
                    for m, dm in enumerate(op.discrete_modes):
                        out[f"{op.name}_e_u{m}"] = np.asarray(dm.e_u_profile, float)
                        out[f"{op.name}_e_v{m}"] = np.asarray(dm.e_v_profile, float)
                        out[f"{op.name}_omega_c{m}"] = np.float64(dm.mode.omega_c)
                        out[f"{op.name}_type{m}"] = np.array(dm.mode.mode_type.value)
                        if dm.mode.z_line is not None:
                            out[f"{op.name}_z_line{m}"] = np.float64(dm.mode.z_line)
                    for k in ("record_scale", "state_scale"):
                        out[f"{op.name}_{k}"] = np.asarray(getattr(op, k), float)
                    out[f"{op.name}_source_scale"] = np.asarray(op._source_scale, float)
                    out[f"{op.name}_dtbc"] = np.array([[np.nan if v is None else v for v in
                                                        (op._dtbc_r[m], op._dtbc_q[m], op._dtbc_z0[m])]
                                                       for m in range(len(op.discrete_modes))])
                    out[f"{op.name}_mur"] = np.array([t is None for t in op._dtbc])
                    out[f"{op.name}_mur_r"] = np.asarray(op._mur_r, float)
                    if op._comp_r_u is not None:
                        for k in ("r_u", "r_v", "live_u", "live_v"):
                            out[f"{op.name}_comp_{k}"] = np.asarray(getattr(op, f"_comp_{k}"), float)

            def wf(t, lo=f_c):
                return gaussian(t, F_MAX) if lo <= 0.0 else modulated_gaussian(t, F_MAX, lo)

            ops[0].set_excitation(excite, wf)
            rec = PortSignalRecorder(dt=dt, ports=ops)
            sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions={}, ports=ops, recorder=rec,
                                      total_time_steps=SNAPS[0], dt=dt, verbose=False,
                                      backend="numpy", precision=prec,
                                      energy_stop_db=None, port_signal_stop_db=None)
            p = f"{prec}_"
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                for n in SNAPS:
                    sol.total_time_steps = n
                    sol.run()
                    out[f"{p}e{n}"] = np.asarray(sol._fields.e_flat, dtype=np.float64).copy()
                    out[f"{p}h{n}"] = np.asarray(sol._fields.h_flat, dtype=np.float64).copy()
                    if n == CHECKPOINT:
                        out.update(flatten(sol.state_dict(), f"{p}state"))
            out[p + "alpha_e"] = np.asarray(sol._alpha_E, np.float64)
            out[p + "beta_e"] = np.asarray(sol._beta_E, np.float64)
            out[p + "alpha_h"] = np.asarray(sol._alpha_H, np.float64)
            out[p + "beta_h"] = np.asarray(sol._beta_H, np.float64)
            sigs = rec.finalize(n_steps_actual=SNAPS[-1])
            for (lbl, m), (V, I) in sigs.items():
                out[f"{p}V_{lbl}_{m}"] = np.asarray(V.values, float)
                out[f"{p}I_{lbl}_{m}"] = np.asarray(I.values, float)
            f_axis = np.linspace(1.1 * f_c if f_c > 0 else 1e9, 24e9, 41)
            t_axis = np.arange(SNAPS[-1]) * dt
            ref = Signal1D(t=t_axis, values=np.array([wf(float(t)) for t in t_axis]), dt=dt,
                           label="excitation")
            s = compute_s_parameters(
                recorder_signals=sigs,
                port_modes={op.name: [dm.mode for dm in op.discrete_modes] for op in ops},
                excited=("port1", excite), reference_signal=ref, f_axis=f_axis,
                port_normal_dx={op.name: op.plane.normal_dx for op in ops},
                port_line_params={(op.name, m): lp for op in ops
                                  for m, lp in op.dtbc_line_params.items()})
            out["f_axis"] = f_axis
            for (lbl, m), v in s.items():
                out[f"{p}S_{lbl}_{m}"] = np.asarray(v, complex)
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(f"wrote {name}.npz ({len(out)} arrays)")


if __name__ == "__main__":
    # python make_c1_fixtures.py [reference src] [case ...]
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src", sys.argv[2:] or None)
