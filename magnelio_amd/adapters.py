"""Adapters from the reference's plug-in objects to this backend's.

A caller holding the reference's objects — its ``Mesh``, ``CPMLBoundary``,
``PlaneWaveSource``, ``PortOperatorModal`` ports, ``MonitorFieldFrequency``
/ ``MonitorFarField`` monitors and the ``FITTimeDomainSolver`` constructor
fields — builds the equivalent HIP solver here; the objects are read
through their attributes only (duck-typed, nothing of the reference is
imported), so the reference's meshes, ports and results stay what they are.

* ``face_key``            BoxFace (src/magnelio/mesh/faces.py:13-30) or its
                          value "z_min" -> "zmin"
* ``mesh_from_reference`` Mesh (src/magnelio/mesh/mesher.py:157-184): grid
                          lines, material ids and library, the PEC edge mask
                          as given (hand edits and thin wires included)
* ``port_from_operator``  PortOperatorModal (ports/_modal/operator.py:505-620):
                          plane maps, calibrated mode profiles, mass slices,
                          V/I scales, Mur and DTBC chain parameters — the
                          operator's own values, not rebuilt
* ``cpml_from_reference`` CPMLBoundary (boundaries/cpml.py:36-90)
* ``plane_wave_from_reference`` PlaneWaveSource (sources/plane_wave.py:65-101)
* ``monitor_from_reference``    MonitorFieldFrequency / MonitorFarField
                          (monitors/field_frequency.py:116-121,
                          far_field.py:98-100)
* ``solver_from_reference``     FITTimeDomainSolver fields (solver/fit_td.py:
                          135-236) -> this package's FITTimeDomainSolver, the
                          K field sets given as (port name, mode) excitations;
                          ``mass_matrices`` may pass the reference's own
                          build_M_eps / build_M_mu / build_M_sigma(_m) vectors
                          (conformal cells, thin wires, donors), which the
                          solver then uses instead of its own cell-category
                          build
* SIBC walls            the reference's SIBCSpec passes to the solver as
                          it is (``sibc=``)
* ``recorder_signals``  the solver's recorded V/I as PortSignalRecorder.
                          finalize returns them (ports/recorder.py:286-340)
"""

from __future__ import annotations

import math

import numpy as np

from . import mesh as _mesh
from .boundaries import CPMLBoundary, PECBoundary, PMCBoundary
from .monitors import MonitorFarField, MonitorFieldFrequency
from .ports import DiscreteMode, PortModel, PortPlane
from .sources import PlaneWaveSource

_FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def face_key(face) -> str:
    """"zmin" for BoxFace.Z_MIN, its value "z_min", or "zmin"."""
    v = getattr(face, "value", face)
    key = str(v).replace("_", "").lower()
    if key not in _FACES:
        raise ValueError(f"not a bbox face: {face!r}")
    return key


def _triple(v) -> tuple:
    a = np.broadcast_to(np.asarray(v, dtype=float), (3,))
    return tuple(float(x) for x in a)


def material_from_reference(mat) -> _mesh.Material:
    """Material (src/magnelio/materials/material.py:18-145); a dispersive one
    carries its DispersionModel's pole-residue pairs (materials/dispersion.py:
    40-103) — ``dispersion`` for eps(w), ``dispersion_mu`` for mu(w) (whose
    high-frequency limit is the material's ``mu``, material.py:106-117)."""
    def poles_of(model):
        return tuple((complex(a), complex(r)) for a, r in model.poles) if model is not None else ()

    return _mesh.Material(name=str(mat.name), epsilon=_triple(mat.epsilon), mu=_triple(mat.mu),
                          sigma=_triple(getattr(mat, "sigma", 0.0)),
                          sigma_m=_triple(getattr(mat, "sigma_m", 0.0)),
                          is_pec=bool(getattr(mat, "is_pec", False)),
                          poles=poles_of(getattr(mat, "dispersion", None)),
                          poles_mu=poles_of(getattr(mat, "dispersion_mu", None)))


def _pec_faces(bc) -> tuple:
    """Faces a reference BoundaryConditions / dict closes with PEC."""
    if bc is None:
        return _FACES  # the reference closes every face with PEC by default
    out = []
    for f in _FACES:
        kind = bc.get(f) if isinstance(bc, dict) else getattr(bc, f, None)
        kind = getattr(kind, "kind", kind)
        if str(getattr(kind, "value", kind)).upper() == "PEC":
            out.append(f)
    return tuple(out)


def mesh_from_reference(mesh) -> _mesh.Mesh:
    g = mesh.grid
    grid = _mesh.GridLines(np.asarray(g.x, float), np.asarray(g.y, float), np.asarray(g.z, float))
    lib = {int(k): material_from_reference(m) for k, m in mesh.material_library.items()}
    out = _mesh.Mesh(grid=grid, material_id=np.ascontiguousarray(mesh.material_id, dtype=np.int32),
                     material_library=lib,
                     pec_mask_edges=np.array(mesh.pec_mask_edges, dtype=bool, copy=True),
                     pec_faces=_pec_faces(getattr(mesh, "boundary_conditions", None)))
    out.pec_rule = False  # the mask is the reference's, whatever edits it carries
    return out


def _opt(v):
    return None if v is None or (isinstance(v, float) and math.isnan(v)) else float(v)


def port_from_operator(op, f_calc: float | None = None) -> PortModel:
    """This backend's port model from a built PortOperatorModal."""
    pl = op.plane
    plane = PortPlane(face=face_key(pl.face), coordinate=float(pl.coordinate),
                      e_u_indices=np.asarray(pl.e_u_indices, np.int64),
                      h_v_indices=np.asarray(pl.h_v_indices, np.int64),
                      u_edge_uv=np.asarray(pl.u_edge_uv, float),
                      u_edge_lengths=np.asarray(pl.u_edge_lengths, float),
                      e_v_indices=np.asarray(pl.e_v_indices, np.int64),
                      h_u_indices=np.asarray(pl.h_u_indices, np.int64),
                      v_edge_uv=np.asarray(pl.v_edge_uv, float),
                      v_edge_lengths=np.asarray(pl.v_edge_lengths, float),
                      e_u_indices_interior=np.asarray(pl.e_u_indices_interior, np.int64),
                      e_v_indices_interior=np.asarray(pl.e_v_indices_interior, np.int64),
                      normal_dx=float(pl.normal_dx),
                      u_node_window=tuple(int(v) for v in pl.u_node_window),
                      v_node_window=tuple(int(v) for v in pl.v_node_window))
    modes = []
    for dm in op.discrete_modes:
        md = dm.mode
        kind = str(getattr(md.mode_type, "value", md.mode_type))
        if kind not in ("TE", "TM", "TEM"):
            raise NotImplementedError(f"port {op.name!r}: {kind} modes are not served by this backend")
        z_line = getattr(md, "z_line", None)
        modes.append(DiscreteMode(name=str(md.name), omega_c=float(md.omega_c), epsilon_r=float(md.epsilon_r),
                            e_u=np.asarray(dm.e_u_profile, float), e_v=np.asarray(dm.e_v_profile, float),
                            h_u=np.asarray(dm.h_u_profile, float), h_v=np.asarray(dm.h_v_profile, float),
                            mode_type=kind, z_line=None if z_line is None else float(z_line)))
    comp_r = comp_live = None
    if getattr(op, "_comp_r_u", None) is not None:  # the complement absorber (operator.py:596-619)
        comp_r = np.concatenate([np.asarray(op._comp_r_u, float), np.asarray(op._comp_r_v, float)])
        comp_live = np.concatenate([np.asarray(op._comp_live_u, float), np.asarray(op._comp_live_v, float)])
    return PortModel(name=str(op.name), plane=plane, modes=modes, dt=float(op._dt),
                     omega_calc=2 * math.pi * f_calc if f_calc else float("nan"),
                     me_u_port=np.asarray(op._me_u_port, float), me_v_port=np.asarray(op._me_v_port, float),
                     me_u_int=np.asarray(op._me_u_int, float), me_v_int=np.asarray(op._me_v_int, float),
                     mh_u=np.asarray(op._mh_u, float), mh_v=np.asarray(op._mh_v, float),
                     record_scale=np.asarray(op.record_scale, float).copy(),
                     source_scale=np.asarray(op._source_scale, float).copy(),
                     state_scale=np.asarray(op.state_scale, float).copy(),
                     dtbc_r=[_opt(v) for v in op._dtbc_r], dtbc_q=[_opt(v) for v in op._dtbc_q],
                     dtbc_z0=[_opt(v) for v in op._dtbc_z0],
                     pair_spread=[_opt(v) for v in op._dtbc_pair_spread],
                     mur_r=np.asarray(op._mur_r, float).copy(), tau=np.asarray(op._tau_m, float).copy(),
                     slab_defect=float(getattr(op, "_chain_slab_defect", 0.0) or 0.0),
                     comp_r=comp_r, comp_live=comp_live)


def cpml_from_reference(bc, grid) -> CPMLBoundary:
    return CPMLBoundary(face_key(bc.face), grid, thickness_cells=int(bc.thickness_cells), m=int(bc.m),
                        R_target=float(bc.R_target), kappa_max=float(bc.kappa_max),
                        alpha_max=float(bc.alpha_max))


def boundary_from_reference(face, bc, grid):
    kind = type(bc).__name__
    if hasattr(bc, "thickness_cells") and hasattr(bc, "R_target"):
        return cpml_from_reference(bc, grid)
    if "PMC" in kind:
        return PMCBoundary(face_key(face))
    if "PEC" in kind:
        return PECBoundary(face_key(face))
    raise TypeError(f"boundary {kind} on {face} is not served by this backend")


def plane_wave_from_reference(src) -> PlaneWaveSource:
    return PlaneWaveSource(direction=tuple(float(v) for v in src.direction),
                           polarization=tuple(float(v) for v in src.polarization),
                           corners=src.corners, amplitude=float(src.amplitude),
                           waveform=str(src.waveform), f_center=src.f_center, f_max=src.f_max)


def monitor_from_reference(mon, **kw):
    """The DFT monitors; ``kw`` passes this backend's block / pipelining options."""
    if hasattr(mon, "margin_cells"):
        return MonitorFarField(np.asarray(mon.freqs, float), margin_cells=int(mon.margin_cells),
                               name=str(mon.name), **kw)
    if hasattr(mon, "fields") and hasattr(mon, "corners"):
        if getattr(mon, "interval", None):
            raise NotImplementedError(f"monitor {mon.name!r}: decimated sampling (interval) is not served")
        return MonitorFieldFrequency(np.asarray(mon.freqs, float), corners=mon.corners,
                                     fields=tuple(mon.fields), name=str(mon.name), **kw)
    raise TypeError(f"monitor {type(mon).__name__} is not served by this backend")


def solver_from_reference(*, mesh, boundary_conditions=None, ports=(), sources=(), monitors=(),
                          total_time_steps=None, dt, energy_stop_db=None, port_signal_stop_db=None,
                          port_signal_min_steps=None, max_time_steps=None, energy_check_interval=None,
                          precision=None, excitations=None, f_calc=None, mass_matrices=None,
                          monitor_options=None, sibc=None, **solver_kw):
    """This backend's FITTimeDomainSolver from the reference solver's fields.

    ``excitations``: per field set a (port name, mode, waveform) — by default
    the excitation set on the reference ports (one set).  ``sibc``: the
    reference's SIBCSpec (_sibc.py:65-89), taken as it is — its surfaces'
    comp / flat_idx / weight / inv_l_dual and its fits' c0 / branches."""
    from .solver import ChannelExcitation, FITTimeDomainSolver  # noqa: PLC0415

    pm = mesh_from_reference(mesh)
    bcs = {face_key(f): boundary_from_reference(f, bc, pm.grid)
           for f, bc in (boundary_conditions or {}).items()}
    hip_ports = [port_from_operator(op, f_calc) for op in ports]
    if excitations is None:
        excitations = []
        for op in ports:
            m = getattr(op, "_excitation_mode", None)
            if m is not None:
                wf = op._excitation_waveform  # already scaled by source_scale
                scale = float(op._source_scale[m])
                excitations.append((op.name, m, lambda t, _f=wf, _s=scale: _f(t) / _s))
        excitations = excitations or [None]
    exc = [None if e is None else ChannelExcitation(e[0], int(e[1]), e[2]) for e in excitations]
    return FITTimeDomainSolver(
        mesh=pm, boundary_conditions=bcs, ports=hip_ports,
        sources=[plane_wave_from_reference(s) for s in sources],
        monitors=[monitor_from_reference(m, **(monitor_options or {})) for m in monitors],
        total_time_steps=total_time_steps, dt=float(dt), energy_stop_db=energy_stop_db,
        port_signal_stop_db=port_signal_stop_db, port_signal_min_steps=port_signal_min_steps,
        max_time_steps=max_time_steps, energy_check_interval=energy_check_interval,
        precision=precision or "single", nsets=len(exc), excitations=exc,
        mass_matrices=mass_matrices, sibc=sibc, **solver_kw)


def recorder_signals(solver, set_index: int = 0, signal_type=None) -> dict:
    """Field set ``set_index``'s recorded V/I in the shape the reference's
    ``PortSignalRecorder.finalize`` returns (ports/recorder.py:286-340):
    ``{(port, mode): (V, I)}`` over ``t = arange(N) * dt``, with the
    per-mode record scale already applied.  ``signal_type`` — the
    reference's ``Signal1D`` — wraps each series as
    ``signal_type(t=, values=, dt=, label=)`` with the recorder's labels
    ``"{port}_mode{m}_V"`` / ``"_I"``; without it the pairs are plain
    float64 arrays."""
    rec = solver.recorded(set_index)
    if signal_type is None:
        return rec
    dt = float(solver.dt)
    out = {}
    for (name, m), (V, I) in rec.items():
        t = np.arange(V.size) * dt
        out[(name, m)] = (signal_type(t=t, values=V, dt=dt, label=f"{name}_mode{m}_V"),
                          signal_type(t=t, values=I, dt=dt, label=f"{name}_mode{m}_I"))
    return out
