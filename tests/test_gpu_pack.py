"""Packed stepping of small solvers (pack.PackedGroup) against the same
solvers stepped alone.

Five seeded jobs of the serving workload (WR-90 guides and crosses of two,
dielectric bricks/spheres, 1-4 modal ports x 1-3 modes, every channel a
field set) are set up twice: once stepped one solver at a time with the
solver's own launches, once as one packed group whose steps are one E-curl
launch over all five grids, the port steps of all their ports in shared
launches, one H-curl launch and one counter launch.  Fields and recorded
port signals are equal bit for bit, in float32 and float64, through eager
steps, a captured graph replayed several times and the eager tail; a member
leaves the group (the group rebuilt without it) and the rest continue equal.
Solvers whose step has more than these launches are refused.
"""

import numpy as np
import pytest
import torch

import bench
from magnelio_amd import pack
from magnelio_amd.sweep import ScatteringSweep

pytestmark = pytest.mark.gpu


def _solvers(jobs, precision):
    out = []
    for j in jobs:
        mesh, ports, dt = j.build()
        sw = ScatteringSweep(mesh, ports, dt=dt, n_steps=j.n_steps, f_axis=j.f_axis, f_max=j.f_max,
                             precision=precision)
        s = sw.make_solver(sw.excited)
        s.setup()
        out.append(s)
    return out


def _state(s):
    torch.cuda.synchronize()
    return (s.e.cpu().numpy().copy(), s.h.cpu().numpy().copy(),
            {k: s.recorded(k) for k in range(s.nsets)})


def _equal(a, b):
    if not (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])):
        return False
    for k in a[2]:
        for ch in a[2][k]:
            va, ia = a[2][k][ch]
            vb, ib = b[2][k][ch]
            if not (np.array_equal(va, vb) and np.array_equal(ia, ib)):
                return False
    return True


@pytest.mark.parametrize("precision", ["single", "double"])
def test_packed_equals_alone(precision, gpu):
    jobs = bench.c5_jobs(n_jobs=5, seed=11, rate=20.0, cells=(16 ** 3, 40 ** 3), steps=(300, 700))
    alone = _solvers(jobs, precision)
    print(f"packed group ({precision}): ports per job {[len(s.ports) for s in alone]}, "
          f"sets {[s.nsets for s in alone]}", flush=True)
    packed = _solvers(jobs, precision)
    plan = (5, 90, 37, 120)  # eager start, a graph capture and replays, eager tails
    steps = sum(plan)
    assert steps <= min(j.n_steps for j in jobs)
    for s in alone:
        s.run(steps)
    grp = pack.PackedGroup(packed)
    for n in plan:
        assert grp.launch(n) == n
    assert grp._graph is not None
    for a, p in zip(alone, packed):
        assert p.n_completed == a.n_completed == steps
        assert _equal(_state(a), _state(p))
    # the first member leaves; the others go on as a rebuilt group
    rest = pack.PackedGroup(packed[1:])
    rest.launch(60)
    for s in alone[1:]:
        s.run(60)
    for a, p in zip(alone[1:], packed[1:]):
        assert _equal(_state(a), _state(p))


def test_refuses_what_it_cannot_pack(gpu):
    from magnelio_amd import mesh as mm
    from magnelio_amd.boundaries import CPMLBoundary
    from magnelio_amd.solver import FITTimeDomainSolver

    mesh = bench.c2_mesh(12, length=12e-3)
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=2) for f in mm.FACES}
    s = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=4,
                            dt=mm.courant_dt(mesh.grid, "normal"), record=False)
    s.setup()
    assert not pack.packable(s)
    with pytest.raises(ValueError, match="cannot be packed"):
        pack.PackedGroup([s])
