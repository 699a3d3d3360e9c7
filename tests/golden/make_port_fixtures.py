"""Write the golden vectors of the modal-port calibration from the reference.

Run in the build container only (the reference never travels to the GPU
machine; only the .npz files this writes do):

    python tests/golden/make_port_fixtures.py /root/reference/src [case ...]

Pins what no closed form fixes: the numerical V/I calibration
(``record_scale``, ``source_scale``, ``state_scale``;
src/magnelio/ports/_modal/operator.py:778-941), the DTBC chain parameters
(r, q, z0; ``_chain_params`` / ``_chain_z0``, operator.py:671-776) and the
M_eps-normalised numerical mode profiles (``build_modal_port``,
ports/_modal/factory.py:1367) on

* ``wr90``:   a WR-90 guide on a uniform 12 x 6 cross-section, 20 cells long;
* ``graded``: the same guide on a seeded non-uniform cross-section (adjacent
  spacing ratio <= 1.3), uniform along z;
* ``corner``: the uniform guide on 12 x 8 with a PEC block over a 3 x 2-cell
  corner of the cross-section along its whole length (an L-shaped section,
  so TE_mn and TM_mn split), 4-mode ports whose channel 3 is the lowest TM
  mode;

all with PEC x/y walls and numerical ports (lowest cut-offs over TE and TM;
3 modes, 4 on ``corner``) on z_min and z_max, dt = courant_dt(grid,
"normal"), f_calc = 25 GHz.  Also stored, under ``op_``: every attribute of the built ``PortOperatorModal``
that a backend reads to run the port (plane maps and geometry, calibrated
mode profiles, mass slices, Mur and DTBC chain parameters) — the adapter
tests build reference-shaped stand-ins from them.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
A, B, L = 22.86e-3, 10.16e-3, 20e-3


def graded_axis(n, length, seed):
    rng = np.random.default_rng(seed)
    w = np.exp(np.concatenate([[0.0], np.cumsum(rng.uniform(np.log(1 / 1.3), np.log(1.3), n - 1))]))
    w = w / w.sum() * length
    return np.concatenate([[0.0], np.cumsum(w)])


# the corner case's PEC block (x0, y0, z0, x1, y1, z1): cells x 0-2, y 0-1
CORNER = (0.0, 0.0, 0.0, 0.25 * A, 0.25 * B, L)


def cases():
    """name -> (x, y, z, corner block?, modes per port)."""
    return {
        "wr90": (np.linspace(0.0, A, 13), np.linspace(0.0, B, 7), np.linspace(0.0, L, 21), False, 3),
        "graded": (graded_axis(12, A, 11), graded_axis(6, B, 12), np.linspace(0.0, L, 21), False, 3),
        "corner": (np.linspace(0.0, A, 13), np.linspace(0.0, B, 9), np.linspace(0.0, L, 21), True, 4),
    }


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    from magnelio._operators.material_matrices import build_M_eps, build_M_mu
    from magnelio.materials.material import Material
    from magnelio.mesh.grid import GridLines
    from magnelio.mesh.mesher import Mesh
    from magnelio.ports._modal import BoxFace, PortSpecNumerical, build_modal_port
    from magnelio.solver.stability import courant_dt

    only = sys.argv[2:]
    for name, (x, y, z, corner, n_modes) in cases().items():
        if only and name not in only:
            continue
        grid = GridLines(x=x, y=y, z=z)
        regions = [(Material.pec(), CORNER)] if corner else []
        mesh = Mesh.from_grid(grid, regions=regions).with_boundary_conditions(
            {"xmin": "PEC", "xmax": "PEC", "ymin": "PEC", "ymax": "PEC", "zmin": "PMC", "zmax": "PMC"})
        dt = courant_dt(grid, "normal")
        m_eps, m_mu = build_M_eps(mesh), build_M_mu(mesh)
        out = {"x": x, "y": y, "z": z, "dt": np.float64(dt), "f_calc": np.float64(25e9),
               "n_modes": np.int64(n_modes), "corner": np.bool_(corner)}
        for label, face in (("port1", BoxFace.Z_MIN), ("port2", BoxFace.Z_MAX)):
            op = build_modal_port(PortSpecNumerical(name=label, plane=face, n_modes=n_modes), mesh,
                                  m_eps, m_mu, dt=dt, f_calc=25e9)
            M = len(op.discrete_modes)
            lp = op.dtbc_line_params
            p = f"{label}_"
            out[p + "record_scale"] = np.asarray(op.record_scale, float)
            out[p + "source_scale"] = np.asarray(op._source_scale, float)
            out[p + "state_scale"] = np.asarray(op.state_scale, float)
            out[p + "omega_c"] = np.array([dm.mode.omega_c for dm in op.discrete_modes])
            out[p + "types"] = np.array([dm.mode.mode_type.value for dm in op.discrete_modes])
            out[p + "dtbc"] = np.array([[lp[m][0], lp[m][1], np.nan if lp[m][2] is None else lp[m][2]]
                                        if m in lp else [np.nan] * 3 for m in range(M)])
            out[p + "kinds"] = np.array(op.termination_kinds)
            for m, dm in enumerate(op.discrete_modes):
                for comp in ("e_u", "e_v", "h_u", "h_v"):
                    out[f"{p}{comp}{m}"] = np.asarray(getattr(dm, comp + "_profile"), float)
            pl = op.plane
            out[p + "e_u_indices"] = np.asarray(pl.e_u_indices, np.int64)
            out[p + "e_v_indices"] = np.asarray(pl.e_v_indices, np.int64)
            o = p + "op_"
            out[o + "face"] = np.array(pl.face.value)
            for k in ("coordinate", "normal_dx"):
                out[o + "plane_" + k] = np.float64(getattr(pl, k))
            for k in ("e_u_indices", "h_v_indices", "u_edge_uv", "u_edge_lengths", "e_v_indices",
                      "h_u_indices", "v_edge_uv", "v_edge_lengths", "e_u_indices_interior",
                      "e_v_indices_interior", "u_node_window", "v_node_window"):
                out[o + "plane_" + k] = np.asarray(getattr(pl, k))
            for k in ("_me_u_port", "_me_v_port", "_me_u_int", "_me_v_int", "_mh_u", "_mh_v",
                      "_mur_r", "_tau_m"):
                out[o + k] = np.asarray(getattr(op, k), float)
            for k in ("_dtbc_r", "_dtbc_q", "_dtbc_z0", "_dtbc_pair_spread"):
                out[o + k] = np.array([np.nan if v is None else v for v in getattr(op, k)], float)
            out[o + "_dt"] = np.float64(op._dt)
            out[o + "_chain_slab_defect"] = np.float64(op._chain_slab_defect or 0.0)
            for m, dm in enumerate(op.discrete_modes):
                out[f"{o}mode{m}_name"] = np.array(dm.mode.name)
                out[f"{o}mode{m}_type"] = np.array(dm.mode.mode_type.value)
                out[f"{o}mode{m}_omega_c"] = np.float64(dm.mode.omega_c)
                out[f"{o}mode{m}_epsilon_r"] = np.float64(dm.mode.epsilon_r)
        np.savez_compressed(HERE / f"ports_{name}.npz", **out)
        print(f"wrote ports_{name}.npz: " + ", ".join(f"{k}" for k in out if k.startswith("port1_")
                                                       and not k[-1].isdigit()))


if __name__ == "__main__":
    # python make_port_fixtures.py [reference src] [case ...]
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
