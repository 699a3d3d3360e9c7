"""GPU checks of kchunk="auto": the tuner times the curl launch shapes on
the solver's zeroed buffers without changing them, caches its choice, a
second solver of the same configuration reads the cache, and the tuned
launches give the default launches' fields bit for bit — as does every shape
the tuner can pick (z-chunk, sets per thread, XCD band)."""

import numpy as np
import pytest
import torch

from magnelio_amd import autotune, kernels
from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary
from magnelio_amd.solver import FITTimeDomainSolver

pytestmark = pytest.mark.gpu


def _solver(kchunk):
    from bench import c2_mesh

    mesh = c2_mesh(24, length=24e-3)
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=4) for f in mm.FACES}
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=30,
                              dt=mm.courant_dt(mesh.grid, "normal"), record=False, nsets=2,
                              kchunk=kchunk)
    sol.setup()
    return sol


def _seed(sol):
    g = torch.Generator(device=sol.e.device).manual_seed(3)
    sol.e.normal_(generator=g).mul_((sol.beta_e != 0).to(sol.e.dtype))


def test_auto_kchunk_tunes_caches_and_matches(tmp_path, monkeypatch, gpu):
    path = tmp_path / "autotune.json"
    monkeypatch.setenv("MAGNELIO_HIP_AUTOTUNE_CACHE", str(path))
    print("autotune: first solver (times the candidates)", flush=True)
    a = _solver("auto")
    assert a.kchunk in autotune.candidates(2)
    assert not bool(a.e.any()) and not bool(a.h.any())  # timing left the zero state alone
    data = autotune._load(path)
    (entry,) = data.values()
    assert entry["shape"] == a.kchunk and len(entry["ms"]) == len(autotune.candidates(2))

    print("autotune: second solver (reads the cache)", flush=True)

    def no_timing(*args, **kwargs):
        raise AssertionError("cached configuration timed again")

    monkeypatch.setattr(autotune, "time_curl", no_timing)
    b = _solver("auto")
    assert b.kchunk == a.kchunk

    print("autotune: tuned vs default launches", flush=True)
    ref = _solver(2)
    for s in (a, ref):
        _seed(s)
        s.run()
    torch.cuda.synchronize()
    assert np.array_equal(a.e.cpu().numpy(), ref.e.cpu().numpy())
    assert np.array_equal(a.h.cpu().numpy(), ref.h.cpu().numpy())


@pytest.mark.parametrize("nsets", [1, 3, 6])
def test_every_shape_gives_the_same_fields(nsets, gpu):
    """Each candidate shape (K = 1, 3, 6; CPML on six faces) against the
    library's default, 12 steps from seeded fields: bit for bit."""
    from bench import c2_mesh

    mesh = c2_mesh(20, length=20e-3)
    bcs = {f: CPMLBoundary(f, mesh.grid, thickness_cells=4) for f in mm.FACES}
    dt = mm.courant_dt(mesh.grid, "normal")

    def run(shape):
        sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, total_time_steps=12, dt=dt,
                                  record=False, nsets=nsets, kchunk=shape)
        sol.setup()
        _seed(sol)
        sol.run()
        torch.cuda.synchronize()
        return sol.e.cpu().numpy(), sol.h.cpu().numpy()

    e0, h0 = run(2)
    for w in autotune.candidates(nsets):
        e, h = run(w)
        assert np.array_equal(e, e0) and np.array_equal(h, h0), kernels.shape_parts(w)
