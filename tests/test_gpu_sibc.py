"""GPU parity of the surface-impedance walls (src/magnelio/solver/_sibc.py,
fit_td.py:419-429, :839-840, :909-916) against the reference's own run of
a graded cavity with a lossy-metal brick, a PEC brick and six PEC walls
(tests/golden/sibc_cavity.npz, the reference's surfaces and ladders):

* float32 fields equal the reference's float32 run bit for bit after 1,
  100 and 300 steps; float64 equal the oracle bit for bit and the
  reference's within 1e-12;
* the branch states of the HIP checkpoint after step 100 carry the
  reference's "sibc" values; the reference's step-100 checkpoint loaded
  into a fresh solver runs on to the reference's step-300 fields;
* a zero impedance leaves the run bit-identical to plain PEC walls;
* K = 2 field sets with graphs: each set equals its single-set run.
"""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import cases
from magnelio_amd import layout
from magnelio_amd import mesh as mm
from magnelio_amd import sibc as ps
from magnelio_amd.boundaries import PECBoundary
from magnelio_amd.solver import FITTimeDomainSolver
from oracle import sibc as osb
from test_sibc_host import _mesh

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from _flat import unflatten  # noqa: E402

pytestmark = pytest.mark.gpu


def _spec(zero=False):
    _, surfs, fits = cases.golden_sibc()
    if zero:
        fits = {t: osb.Fit(0.0) for t in fits}
    return ps.SIBCSpec(tuple(surfs), fits)


def _e0(f, seed=11):
    e = np.random.default_rng(seed).standard_normal(f["pec_edges"].size)
    e[f["pec_edges"]] = 0.0
    return e


def _solver(precision, steps, spec, nsets=1, use_graphs=False):
    f, mesh = _mesh()
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions={fc: PECBoundary(fc) for fc in ps.FACES},
                              total_time_steps=steps, dt=float(f["dt"]), precision=precision,
                              sibc=spec, nsets=nsets, use_graphs=use_graphs)
    return f, mesh, sol


def _seed(sol, f, s, seed, dtype):
    sol.e[s].copy_(torch.from_numpy(layout.to_device(_e0(f, seed).astype(dtype), *sol.dims, "E")))


@pytest.mark.parametrize("precision,dtype", [("single", np.float32), ("double", np.float64)])
def test_sibc_matches_the_reference(precision, dtype, gpu):
    print(f"SIBC cavity {precision}", flush=True)
    f, mesh, sol = _solver(precision, 300, _spec())
    sol.setup()
    assert sol._sibc is not None and len(sol._sibc.blocks) == 8
    _seed(sol, f, 0, 11, dtype)
    orc = None
    if dtype == np.float64:
        orc, _, _ = cases.golden_sibc_oracle(np.float64)
    p = precision + "_"
    for n in (1, 100, 300):
        sol.run(n - sol.n_completed)
        e, h = sol.fields(0)
        if orc is not None:
            orc.run(n - orc.n)
        for comp, got in (("e", e), ("h", h)):
            want = f[f"{p}{comp}{n}"]
            err = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
            print(f"  step {n} {comp}: rel {err:.2e}", flush=True)
            if dtype == np.float32:
                assert np.array_equal(got.astype(np.float64), want), (n, comp)
            else:
                assert err <= 1e-12, (n, comp)
                assert np.array_equal(got, orc.e if comp == "e" else orc.h), (n, comp)
        if n == 100:
            sd = sol.state_dict()
            for tag, d in sd["sibc"].items():
                for k, u in d.items():
                    ref = f[f"{p}state/sibc/{tag}/{k}"]
                    assert np.max(np.abs(u - ref)) <= 1e-12 * max(np.abs(ref).max(), 1e-300), (tag, k)


def test_reference_checkpoint_resumes(gpu):
    print("SIBC cavity: the reference's step-100 checkpoint -> step 300", flush=True)
    f, mesh, sol = _solver("double", 300, _spec())
    ref_sd = unflatten({k: f[k] for k in f if k.startswith("double_state/")}, "double_state")
    sol.load_state_dict(ref_sd)
    assert sol.n_completed == 100
    sol.run()
    e, h = sol.fields(0)
    for comp, got in (("e", e), ("h", h)):
        want = f[f"double_{comp}300"]
        assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want), comp


def test_zero_impedance_is_plain_pec(gpu):
    print("SIBC cavity: zero impedance == PEC walls", flush=True)
    out = []
    for spec in (_spec(zero=True), None):
        f, mesh, sol = _solver("single", 120, spec)
        sol.setup()
        assert sol._sibc is None
        _seed(sol, f, 0, 11, np.float32)
        sol.run()
        out.append(sol.fields(0))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_sibc_sets_and_graphs(gpu):
    print("SIBC cavity: K=2 with graphs vs single sets", flush=True)
    one = []
    for seed in (11, 4):
        f, mesh, sol = _solver("single", 160, _spec())
        sol.setup()
        _seed(sol, f, 0, seed, np.float32)
        sol.run()
        one.append(sol.fields(0))
    f, mesh, two = _solver("single", 160, _spec(), nsets=2, use_graphs=True)
    two.setup()
    _seed(two, f, 0, 11, np.float32)
    _seed(two, f, 1, 4, np.float32)
    two.run()
    for s in range(2):
        e, h = two.fields(s)
        assert np.array_equal(e, one[s][0]) and np.array_equal(h, one[s][1]), s
