"""The precision-study case (C4) against the oracle, at reduced size.

24^3 cells of the C4 layout: a 32 x 20 x 48 mm guide with PEC x/y walls, a
seeded slab of two-pole Lorentz material with sigma 0.05 S/m, a 2-mode DTBC
port on the open end z_min and an 8-layer CPML on the open end z_max (the
ADE completion between the curl and the CPML pass); both port modes excited
as two field sets, 500 steps; the oracle runs each excitation (its ADE,
CPML and port restatements).

* float64: fields within 1e-10, S11 (2 x 2) within 1e-9;
* float32: fields within 1e-5 of the peak field norm, S within 1e-4 — the
  precision study's gates, here with the CPML the configuration places at
  the open end.
"""

import math

import numpy as np
import pytest

import cases
from oracle import sparams as osp

pytestmark = pytest.mark.gpu

N, STEPS = 24, 500
CHANNELS = [("port1", 0), ("port1", 1)]


@pytest.fixture(scope="module")
def c4_oracle_runs():
    out = {}
    for ch in CHANNELS:
        print(f"oracle C4 ({N}^3) float64, excitation {ch}, {STEPS} steps (CPU)", flush=True)
        orc, ports, dt, _ = cases.c4_oracle(N, ch)
        peak = 0.0
        for _ in range(STEPS):
            orc.step()
            peak = max(peak, float(np.linalg.norm(orc.e)))
        f_axis = np.linspace(1.1 * ports[0].modes[1]["wc"] / (2 * math.pi), 19e9, 41)
        sig = {(p.name, m): (np.array(p.V[m]), np.array(p.I[m])) for p in ports for m in range(2)}
        lines = {(p.name, m): p.line[m] for p in ports for m in range(2)}
        out[ch] = dict(e=orc.e.copy(), peak=peak, S=osp.s_params(sig, lines, ch, f_axis, dt),
                       f_axis=f_axis)
    return out


@pytest.mark.parametrize("precision", ["double", "single"])
def test_c4_with_cpml_vs_oracle(c4_oracle_runs, precision, gpu):
    from bench import c4_case
    from magnelio_amd.signals import s_parameters
    from magnelio_amd.sweep import ScatteringSweep

    print(f"HIP C4 ({N}^3) {precision}, 2 sets, {STEPS} steps", flush=True)
    mesh, ports, dt, bcs, _ = c4_case(N)
    sweep = ScatteringSweep(mesh, ports, dt=dt, n_steps=STEPS, f_axis=np.array([1e10]), f_max=20e9,
                            excited=CHANNELS, precision=precision, boundary_conditions=bcs)
    sol = sweep.make_solver(CHANNELS)
    sol.run()
    assert sol.cpml is not None and sol._dispersion is not None
    fgate = 1e-10 if precision == "double" else 1e-5
    sgate = 1e-9 if precision == "double" else 1e-4
    modes = {p.name: p.modes for p in ports}
    lines = {(p.name, m): lp for p in ports for m, lp in p.dtbc_line_params.items()}
    ndx = {p.name: p.plane.normal_dx for p in ports}
    for s, ch in enumerate(CHANNELS):
        ref = c4_oracle_runs[ch]
        e, _ = sol.fields(s)
        ee = float(np.linalg.norm(e.astype(np.float64) - ref["e"]))
        rel = ee / (np.linalg.norm(ref["e"]) if precision == "double" else ref["peak"])
        S = s_parameters(sol.recorded(s), modes, ch, ref["f_axis"], dt, port_normal_dx=ndx,
                         port_line_params=lines)
        ds = max(float(np.nanmax(np.abs(S[k] - ref["S"][k]))) for k in ref["S"])
        print(f"  set {s} {ch}: field {rel:.2e}  |dS| {ds:.2e}", flush=True)
        assert rel <= fgate and ds <= sgate, ch
