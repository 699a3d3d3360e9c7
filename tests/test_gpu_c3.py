"""The batched multi-excitation case (C3) against the oracle, at reduced size.

48 x 24 x 72 cells of the C3 layout: WR-90 guide, two seeded eps_r = 30
posts, a PEC iris with a seeded off-centre window, 3-mode DTBC ports on
z_min / z_max: K = 6 field sets (every port mode excited) in one batched
solver with HIP graphs and a pipelined field-plane E DFT (33 frequencies,
64-step blocks, 4 blocks per period), 600 steps; the oracle runs the six
excitations one at a time (cells filled independently from the seeded
layout, closed-form TE modes on the uniform feed).

* float64: fields of every set within 1e-10, the 6 x 6 S-matrix on 41
  frequencies within 1e-9, the plane DFT within 1e-6 of its peak (float32
  MFMA contraction);
* float32: fields within 1e-5 of the peak field norm, S within 1e-4, DFT
  within 1e-5 of its peak;
* the window's off-centre position couples the modes: cross-mode
  transmissions are non-zero.
"""

import math

import numpy as np
import pytest

import cases
from oracle import monitors as omon
from oracle import sparams as osp

pytestmark = pytest.mark.gpu

STEPS = 600
FREQS = np.linspace(7e9, 24e9, 33)
CHANNELS = [("port1", 0), ("port1", 1), ("port1", 2), ("port2", 0), ("port2", 1), ("port2", 2)]


def _plane(dims):
    nx, ny, nz = dims
    kz = round(0.3 * nz)
    return (slice(0, nx), slice(0, ny), slice(kz, kz + 1))


@pytest.fixture(scope="module")
def c3_oracle_runs():
    out = {}
    for ch in CHANNELS:
        print(f"oracle C3 (48x24x72) float64, excitation {ch}, {STEPS} steps (CPU)", flush=True)
        orc, ports, dt, _ = cases.c3_oracle(ch)
        x, y, z = (np.linspace(0.0, s, n + 1) for s, n in zip(cases.C3_SIZE, cases.C3_SMALL))
        mon = omon.FieldDFT(FREQS, x, y, z, _plane(cases.C3_SMALL), comps=("Ex", "Ey", "Ez"))
        orc.monitors.append(mon)
        peak = 0.0
        for _ in range(STEPS):
            orc.step()
            peak = max(peak, float(np.linalg.norm(orc.e)))
        f_axis = np.linspace(1.1 * ports[0].modes[0]["wc"] / (2 * math.pi), 24e9, 41)
        sig = {(p.name, m): (np.array(p.V[m]), np.array(p.I[m])) for p in ports for m in range(3)}
        lines = {(p.name, m): p.line[m] for p in ports for m in range(3)}
        S = osp.s_params(sig, lines, ch, f_axis, dt)
        out[ch] = dict(e=orc.e.copy(), h=orc.h.copy(), peak=peak, S=S, bins=mon.bins, f_axis=f_axis)
    return out


def _hip(precision):
    from bench import c3_case
    from magnelio_amd.monitors import MonitorFieldFrequency
    from magnelio_amd.sweep import ScatteringSweep

    mesh, ports, dt = c3_case(dims=cases.C3_SMALL)

    def monitors():
        return [MonitorFieldFrequency(FREQS, region=_plane(cases.C3_SMALL), fields=("E",), block=64)]

    sweep = ScatteringSweep(mesh, ports, dt=dt, n_steps=STEPS, f_axis=np.array([1e10]), f_max=25e9,
                            precision=precision, monitors_factory=monitors, graph_block=16)
    sol = sweep.make_solver(CHANNELS)
    sol.run()
    assert sol._graph is not None, getattr(sol, "graph_error", "no graph")
    return sweep, sol


@pytest.mark.parametrize("precision", ["double", "single"])
def test_c3_batched_vs_oracle(c3_oracle_runs, precision, gpu):
    print(f"HIP C3 (48x24x72) {precision}, K = 6 batched, {STEPS} steps", flush=True)
    sweep, sol = _hip(precision)
    mon = sol.monitors[0]
    fgate = 1e-10 if precision == "double" else 1e-5
    sgate = 1e-9 if precision == "double" else 1e-4
    dgate = 1e-6 if precision == "double" else 1e-5
    cross = 0.0
    for s, ch in enumerate(CHANNELS):
        ref = c3_oracle_runs[ch]
        e, h = sol.fields(s)
        ee = float(np.linalg.norm(e.astype(np.float64) - ref["e"]))
        rel = ee / (np.linalg.norm(ref["e"]) if precision == "double" else ref["peak"])
        rec = sol.recorded(s)
        from magnelio_amd.signals import s_parameters

        modes = {p.name: p.modes for p in sweep.ports}
        lines = {(p.name, m): lp for p in sweep.ports for m, lp in p.dtbc_line_params.items()}
        ndx = {p.name: p.plane.normal_dx for p in sweep.ports}
        S = s_parameters(rec, modes, ch, ref["f_axis"], sweep.dt, port_normal_dx=ndx, port_line_params=lines)
        ds = max(float(np.nanmax(np.abs(S[k] - ref["S"][k]))) for k in ref["S"])
        got = mon.result(s)
        peak = max(np.abs(v).max() for v in ref["bins"].values())
        dd = max(float(np.abs(got[c][:, :, :, 0] - ref["bins"][c][..., 0]).max()) for c in ref["bins"]) / peak
        print(f"  set {s} {ch}: field {rel:.2e}  |dS| {ds:.2e}  DFT {dd:.2e}", flush=True)
        assert rel <= fgate and ds <= sgate and dd <= dgate, ch
        if ch[0] == "port1":
            for m in range(3):
                if m != ch[1]:
                    cross = max(cross, float(np.nanmax(np.abs(S[("port2", m)]))))
    print(f"  largest cross-mode transmission {cross:.3f}", flush=True)
    assert cross > 1e-2
