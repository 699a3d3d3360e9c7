"""GPU parity of lumped RLC elements and discrete ports.

A Gaussian-driven discrete port (SeriesRLC(R=Z0)) across the middle of a
PEC box, plus passive series-RLC and parallel-RLC loads on other edge
chains, against the oracle's restatement of LumpedElementOperator:

* FP64 and FP32: fields and recorded V/I bit-identical to the oracle;
* the Thevenin identity V + Z0 I = 2 sqrt(Z0) a(t) holds on the device
  record to 1e-12 (test_lumped_port.py:203-238);
* K = 2 field sets (set 0 driven, set 1 idle): set 0 equals the single run
  bit for bit and set 1 stays exactly zero;
* graph replay equals eager launches.
"""

import math

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import PECBoundary
from magnelio_amd.lumped import (LumpedElement, ParallelRLC, PortSpecLumped, SeriesRLC,
                                 build_lumped_element, build_lumped_port)
from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver
from oracle import lumped as ol
from test_oracle_lumped import gap_case

pytestmark = pytest.mark.gpu

LOADS = [("rlc_s", (6, 2, 2), (6, 2, 7), "series", dict(R=20.0, L=2e-9, C=0.5e-12)),
         ("rlc_p", (3, 7, 5), (9, 7, 5), "parallel", dict(R=80.0, L=1e-9, C=0.2e-12))]


def _oracle(dtype, steps):
    orc, port, a_t, (x, y, z), (start, end), _ = gap_case(dtype, steps=steps)
    loads = []
    for name, p0, p1, kind, kw in LOADS:
        s = (x[p0[0]], y[p0[1]], z[p0[2]])
        e = (x[p1[0]], y[p1[1]], z[p1[2]])
        _, _, flat = ol.snap_chain(x, y, z, s, e)
        elem = ol.SeriesRLC(**kw) if kind == "series" else ol.ParallelRLC(**kw)
        me = _me(x, y, z)
        loads.append(ol.Lumped(name, flat, orc.dt / me[flat], elem, orc.dt, recorded=False))
    orc.lumped += loads
    orc.run(steps)
    return orc, port, a_t, (x, y, z), (start, end)


def _me(x, y, z):
    from oracle import mesh as om
    dims = (x.size - 1, y.size - 1, z.size - 1)
    return om.mass_e(x, y, z, np.zeros(dims, np.int32), [[1.0]] * 3, om.EPS0)


def _hip(axes, start_end, a_t, precision, steps, nsets=1, use_graphs=False):
    x, y, z = axes
    grid = mm.GridLines(x, y, z)
    mesh = mm.Mesh.from_grid(grid)
    dt = mm.courant_dt(grid, "normal")
    me = mm.build_m_eps(mesh)
    ops = [build_lumped_port(PortSpecLumped("p1", *start_end, Z0=50.0), mesh, me, dt)]
    for name, p0, p1, kind, kw in LOADS:
        s = (x[p0[0]], y[p0[1]], z[p0[2]])
        e = (x[p1[0]], y[p1[1]], z[p1[2]])
        elem = SeriesRLC(**kw) if kind == "series" else ParallelRLC(**kw)
        ops.append(build_lumped_element(LumpedElement(name, s, e, elem), mesh, me, dt))
    exc = [ChannelExcitation("p1", 0, a_t)] + [None] * (nsets - 1)
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions={f: PECBoundary(f) for f in mm.FACES},
                              lumped=ops, excitations=exc, nsets=nsets, total_time_steps=steps,
                              dt=dt, precision=precision, use_graphs=use_graphs)
    sol.run()
    return sol


@pytest.mark.parametrize("precision,dtype", [("double", np.float64), ("single", np.float32)])
def test_lumped_bit_exact(precision, dtype, gpu):
    steps = 300
    print(f"lumped gap case {precision}", flush=True)
    orc, port, a_t, axes, se = _oracle(dtype, steps)
    sol = _hip(axes, se, a_t, precision, steps)
    assert sol.dt == orc.dt
    e, h = sol.fields(0)
    assert np.array_equal(e, orc.e)
    assert np.array_equal(h, orc.h)
    V, I = sol.recorded(0)[("p1", 0)]
    assert np.array_equal(V, np.array(port.V))
    assert np.array_equal(I, np.array(port.I))
    src = np.array([2 * math.sqrt(50.0) * a_t(n * orc.dt + orc.dt) for n in range(steps)])
    rel = np.abs(V + 50.0 * I - src).max() / np.abs(src).max()
    print(f"  Thevenin residual {rel:.2e}", flush=True)
    assert rel < 1e-12 if precision == "double" else rel < 1e-6


def test_batched_idle_set_and_graphs(gpu):
    steps = 256
    print("lumped K=2 batch, eager vs graphs", flush=True)
    orc, port, a_t, axes, se = _oracle(np.float32, steps)
    single = _hip(axes, se, a_t, "single", steps)
    batch = _hip(axes, se, a_t, "single", steps, nsets=2)
    graph = _hip(axes, se, a_t, "single", steps, nsets=2, use_graphs=True)
    assert graph._graph is not None, getattr(graph, "graph_error", "graph not captured")
    e1, h1 = single.fields(0)
    for sol in (batch, graph):
        e0, h0 = sol.fields(0)
        assert np.array_equal(e0, e1) and np.array_equal(h0, h1)
        ez, hz = sol.fields(1)
        assert not ez.any() and not hz.any()
        assert np.array_equal(sol.recorded(0)[("p1", 0)][0], single.recorded(0)[("p1", 0)][0])
