"""Thin wire in a square PEC duct on the GPU: the reference's line gates
(test_thin_wire_line.py:1-60, :178-273) on this engine.

A wire along z centred in a 12 x 12-cell duct, 60 cells long, fed through a
lumped gap port two cells in and terminated by a lumped resistor three
cells before the far wall; dt = spectral_dt of the corrected mesh.
Closed form of the square coax: Z0 = (eta0 / 2 pi) ln(1.0787 s / d).

* T1: the median input impedance Re Zin matches Z0 within 5 % for
  a = 0.05, 0.1, 0.2 cells, and Z0(0.05) - Z0(0.2) = (eta0/2pi) ln 4
  within 15 %;
* T3: without the correction the line is radius-blind (identical S11 for
  two radii) and solves to the grid's own equivalent radius r0 / cell in
  [0.15, 0.30].
"""

import math

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd.lumped import PortSpecLumped, SeriesRLC, build_lumped_port
from magnelio_amd.signals import resolve_excitation, s_parameters
from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver
from magnelio_amd.stability import spectral_dt
from magnelio_amd.thinwire import ThinWire, apply_thin_wires

pytestmark = pytest.mark.gpu

ETA0_2PI = 59.9585
D = 1e-3
S_CELLS, L_CELLS, K_FEED, N_STEPS = 12, 60, 2, 6000
F_AXIS = np.linspace(0.5e9, 4.5e9, 9)
_CACHE = {}


def duct_z0(a):
    return ETA0_2PI * math.log(1.0787 * (S_CELLS * D) / (2.0 * a))


def line_run(a, correction=True, port_z0=None, term_r=None):
    key = (a, correction, port_z0, term_r)
    if key in _CACHE:
        return _CACHE[key]
    uni = np.linspace(0, S_CELLS * D, S_CELLS + 1)
    g = mm.GridLines(uni, uni.copy(), np.linspace(0, L_CELLS * D, L_CELLS + 1))
    mesh = mm.Mesh.from_grid(g)
    nz = L_CELLS
    k_term = nz - 3
    x0 = y0 = 6e-3
    legs = [(0, K_FEED), (K_FEED + 1, k_term), (k_term + 1, nz)]
    wires = [ThinWire(f"w{q}", ((x0, y0, g.z[k0]), (x0, y0, g.z[k1])), a) for q, (k0, k1) in enumerate(legs)]
    apply_thin_wires(mesh, wires, correction=correction)
    dt = spectral_dt(mesh, "normal")
    me = mm.build_m_eps(mesh)
    z0p = port_z0 if port_z0 is not None else duct_z0(a)
    r_term = term_r if term_r is not None else duct_z0(a)
    feed = build_lumped_port(PortSpecLumped("feed", (x0, y0, g.z[K_FEED]), (x0, y0, g.z[K_FEED + 1]), Z0=z0p),
                             mesh, me, dt)
    term = build_lumped_port(PortSpecLumped("term", (x0, y0, g.z[k_term]), (x0, y0, g.z[k_term + 1]),
                                            Z0=r_term, element=SeriesRLC(R=r_term)), mesh, me, dt)
    wf = resolve_excitation(0.0, 0.0, 5e9)
    sol = FITTimeDomainSolver(mesh=mesh, lumped=[feed, term], total_time_steps=N_STEPS, dt=dt,
                              excitations=[ChannelExcitation("feed", 0, wf)], use_graphs=True,
                              graph_block=50)
    sol.run()
    s = s_parameters(sol.recorded(0), {}, ("feed", 0), F_AXIS, dt,
                     lumped_z0={"feed": z0p, "term": r_term})
    s11 = s[("feed", 0)]
    zin = z0p * (1 + s11) / (1 - s11)
    _CACHE[key] = (s11, zin)
    return s11, zin


@pytest.mark.parametrize("a_frac", [0.05, 0.1, 0.2])
def test_t1_z0_radius_sweep(a_frac, gpu):
    a = a_frac * D
    print(f"thin-wire line a = {a_frac} D", flush=True)
    s11, zin = line_run(a)
    z_med = float(np.median(zin.real))
    print(f"  median Re Zin {z_med:.1f} vs analytic {duct_z0(a):.1f} ohm, "
          f"max |S11| {20 * np.log10(np.abs(s11).max()):.1f} dB")
    assert abs(z_med - duct_z0(a)) < 0.05 * duct_z0(a)


def test_t1_ln_tracking(gpu):
    _, zin_thin = line_run(0.05 * D)
    _, zin_fat = line_run(0.2 * D)
    dz = float(np.median(zin_thin.real) - np.median(zin_fat.real))
    ref = ETA0_2PI * math.log(4.0)
    assert abs(dz - ref) < 0.15 * ref


def test_t3_bare_line(gpu):
    s_a, zin = line_run(0.05 * D, correction=False, port_z0=200.0, term_r=200.0)
    s_b, _ = line_run(0.2 * D, correction=False, port_z0=200.0, term_r=200.0)
    np.testing.assert_array_equal(s_a, s_b)
    z_bare = float(np.median(zin.real))
    r0 = 1.0787 * (S_CELLS * D) / (2.0 * math.exp(z_bare / ETA0_2PI))
    print(f"bare grid r0 = {r0 / D:.3f} cells")
    assert 0.15 <= r0 / D <= 0.30
    assert abs(z_bare - duct_z0(0.05 * D)) > 50.0
